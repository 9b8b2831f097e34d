"""ctypes wrapper of the CPU oracle (test infrastructure only).

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker; the product library never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class CoarsenOpts(ctypes.Structure):
    _fields_ = [
        ("parts", ctypes.c_int), ("seed", ctypes.c_uint32), ("stop_ratio", ctypes.c_double),
        ("max_levels", ctypes.c_int), ("cluster_size_factor", ctypes.c_int),
        ("pack_isolated", ctypes.c_int), ("progressive", ctypes.c_int),
        ("fine_factor", ctypes.c_int), ("capacity_growth", ctypes.c_int),
        ("seed_matching", ctypes.c_int), ("fast_path", ctypes.c_int),
        ("hem_degree_limit", ctypes.c_int),
    ]


class RefineOpts(ctypes.Structure):
    _fields_ = [
        ("parts", ctypes.c_int), ("imbalance", ctypes.c_double), ("max_rounds", ctypes.c_int),
        ("stale_rounds", ctypes.c_int), ("growth_ratio", ctypes.c_double),
        ("exchange_rounds", ctypes.c_int), ("graph_loop_threshold", ctypes.c_int),
        ("exchange_max_n", ctypes.c_int), ("exchange_only", ctypes.c_int),
    ]


class LevelStats(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64), ("m", ctypes.c_int64), ("refined", ctypes.c_int),
        ("initial_repair_failed", ctypes.c_int), ("rounds", ctypes.c_int),
        ("initial_cut", ctypes.c_uint64), ("final_cut", ctypes.c_uint64),
        ("lp_moves", ctypes.c_uint64), ("repair_moves", ctypes.c_uint64),
        ("group_moves", ctypes.c_uint64), ("repair_iterations", ctypes.c_int),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        _LIB = ctypes.CDLL(path)
        L = _LIB
        vp = ctypes.c_void_p
        L.og_coarsen.restype = vp
        L.og_coarsen.argtypes = [ctypes.c_int64, ctypes.c_int64, vp, vp, vp, vp, vp]
        L.og_hier_levels.argtypes = [vp]
        L.og_hier_level_info.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp, vp]
        L.og_hier_level_copy.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp]
        L.og_hier_map_copy.argtypes = [vp, ctypes.c_int, vp]
        L.og_hier_free.argtypes = [vp]
        L.og_uncoarsen.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.og_refine_graph.argtypes = [ctypes.c_int64, ctypes.c_int64, vp, vp, vp, vp, vp, vp,
                                      ctypes.c_int, ctypes.c_uint64, vp]
        L.og_cut.restype = ctypes.c_uint64
        L.og_cut.argtypes = [ctypes.c_int64, vp, vp, vp, vp]
        L.og_capacity.restype = ctypes.c_uint64
        L.og_capacity.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_double]
        L.og_contract_graph.argtypes = [ctypes.c_int64, ctypes.c_int64, vp, vp, vp, vp, ctypes.c_int64,
                                        vp, vp, vp, vp]
        L.og_last_error.restype = ctypes.c_char_p
        L.og_set_threads.argtypes = [ctypes.c_int]
        L.og_get_threads.restype = ctypes.c_int
        L.og_default_coarsen_opts.argtypes = [ctypes.c_int, ctypes.c_uint32, vp]
        L.og_default_refine_opts.argtypes = [ctypes.c_int, ctypes.c_double, vp]
        L.og_admission_batch.argtypes = [ctypes.c_int64, vp, vp, vp, vp, vp, vp, ctypes.c_uint64, vp]
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _err():
    return lib().og_last_error().decode()


def coarsen_opts(parts, seed=0, **overrides):
    o = CoarsenOpts()
    lib().og_default_coarsen_opts(parts, seed, ctypes.byref(o))
    for key, value in overrides.items():
        setattr(o, key, value)
    return o


def refine_opts(parts, imbalance=1.03, **overrides):
    o = RefineOpts()
    lib().og_default_refine_opts(parts, imbalance, ctypes.byref(o))
    for key, value in overrides.items():
        setattr(o, key, value)
    return o


def set_threads(threads):
    """Threads of the oracle's OpenMP loops (results do not depend on it)."""
    lib().og_set_threads(int(threads))


def get_threads():
    return int(lib().og_get_threads())


def capacity(total, parts, imbalance=1.03):
    return int(lib().og_capacity(total, parts, imbalance))


def cut(off, nbr, ew, labels):
    off = np.ascontiguousarray(off, np.int32)
    nbr = np.ascontiguousarray(nbr, np.int32)
    ew = None if ew is None else np.ascontiguousarray(ew, np.uint32)
    labels = np.ascontiguousarray(labels, np.int32)
    return int(lib().og_cut(len(off) - 1, _p(off), _p(nbr), _p(ew), _p(labels)))


class Hierarchy:
    """Coarsening hierarchy produced by the oracle."""

    def __init__(self, off, nbr, ew, vw, opts):
        self.off = np.ascontiguousarray(off, np.int32)
        self.nbr = np.ascontiguousarray(nbr, np.int32)
        self.ew = None if ew is None else np.ascontiguousarray(ew, np.uint32)
        self.vw = None if vw is None else np.ascontiguousarray(vw, np.uint32)
        self.opts = opts
        n = len(self.off) - 1
        self.h = lib().og_coarsen(n, len(self.nbr), _p(self.off), _p(self.nbr), _p(self.ew), _p(self.vw),
                                  ctypes.byref(opts))
        if not self.h:
            raise RuntimeError("oracle coarsen failed: " + _err())

    def __del__(self):
        if getattr(self, "h", None):
            lib().og_hier_free(self.h)
            self.h = None

    @property
    def levels(self):
        return lib().og_hier_levels(self.h)

    def info(self, level):
        n = ctypes.c_int64(); m = ctypes.c_int64(); md = ctypes.c_uint64(); hh = ctypes.c_int()
        ib = ctypes.c_uint64()
        lib().og_hier_level_info(self.h, level, ctypes.byref(n), ctypes.byref(m), ctypes.byref(md),
                                 ctypes.byref(hh), ctypes.byref(ib))
        return dict(n=n.value, m=m.value, max_degree=md.value, has_heavy=hh.value, incident_bound=ib.value)

    def graph(self, level):
        i = self.info(level)
        off = np.zeros(i["n"] + 1, np.int32); nbr = np.zeros(i["m"], np.int32)
        ew = np.zeros(i["m"], np.uint32); vw = np.zeros(i["n"], np.uint32)
        lib().og_hier_level_copy(self.h, level, _p(off), _p(nbr), _p(ew), _p(vw))
        return off, nbr, ew, vw

    def map(self, level):
        i = self.info(level)
        out = np.zeros(i["n"], np.int32)
        lib().og_hier_map_copy(self.h, level, _p(out))
        return out

    def uncoarsen(self, coarsest_labels, ropts):
        n0 = self.info(0)["n"]
        out = np.zeros(n0, np.int32)
        loads = np.zeros(ropts.parts, np.uint64)
        c = ctypes.c_uint64()
        stats = (LevelStats * self.levels)()
        labels = np.ascontiguousarray(coarsest_labels, np.int32)
        rc = lib().og_uncoarsen(self.h, _p(labels), ctypes.byref(ropts), _p(out), ctypes.byref(c), _p(loads),
                                ctypes.byref(stats))
        if rc != 0:
            raise RuntimeError("oracle uncoarsen failed: " + _err())
        return out, c.value, loads, [dict((f, getattr(s, f)) for f, _ in LevelStats._fields_) for s in stats]


def refine_graph(off, nbr, ew, vw, labels, ropts, known_cut=None):
    off = np.ascontiguousarray(off, np.int32); nbr = np.ascontiguousarray(nbr, np.int32)
    ew = None if ew is None else np.ascontiguousarray(ew, np.uint32)
    vw = None if vw is None else np.ascontiguousarray(vw, np.uint32)
    out = np.array(labels, np.int32, copy=True)
    st = LevelStats()
    rc = lib().og_refine_graph(len(off) - 1, len(nbr), _p(off), _p(nbr), _p(ew), _p(vw), _p(out),
                               ctypes.byref(ropts), int(known_cut is not None), int(known_cut or 0),
                               ctypes.byref(st))
    if rc != 0:
        raise RuntimeError("oracle refine failed: " + _err())
    return out, dict((f, getattr(st, f)) for f, _ in LevelStats._fields_)


def contract(off, nbr, ew, cmap, nc):
    off = np.ascontiguousarray(off, np.int32); nbr = np.ascontiguousarray(nbr, np.int32)
    ew = None if ew is None else np.ascontiguousarray(ew, np.uint32)
    cmap = np.ascontiguousarray(cmap, np.int32)
    coff = np.zeros(nc + 1, np.int32); cnbr = np.zeros(len(nbr), np.int32); cew = np.zeros(len(nbr), np.uint32)
    cm = ctypes.c_int64()
    rc = lib().og_contract_graph(len(off) - 1, len(nbr), _p(off), _p(nbr), _p(ew), _p(cmap), nc,
                                 _p(coff), _p(cnbr), _p(cew), ctypes.byref(cm))
    if rc != 0:
        raise RuntimeError("oracle contract failed: " + _err())
    return coff, cnbr[:cm.value].copy(), cew[:cm.value].copy()


def admission_batch(targets, keys, ties, vertices, weights, bases, cap):
    """One coarsening admission batch: accepted flags in request order."""
    n = len(targets)
    t = np.ascontiguousarray(targets, np.int32); k = np.ascontiguousarray(keys, np.uint64)
    ti = np.ascontiguousarray(ties, np.uint32); v = np.ascontiguousarray(vertices, np.int32)
    w = np.ascontiguousarray(weights, np.uint64); b = np.ascontiguousarray(bases, np.uint64)
    out = np.zeros(max(n, 1), np.uint8)
    lib().og_admission_batch(n, _p(t), _p(k), _p(ti), _p(v), _p(w), _p(b), cap, _p(out))
    return out[:n].astype(bool)


# ------------------------------------------------------------ single stages
_STAGES_BOUND = False


def _bind_stages():
    global _STAGES_BOUND
    if _STAGES_BOUND:
        return
    L = lib()
    vp, i64, u64, u32, ci, dbl = (ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64, ctypes.c_uint32,
                                  ctypes.c_int, ctypes.c_double)
    L.og_t_seed_match.restype = i64
    L.og_t_seed_match.argtypes = [i64, vp, vp, vp, vp, u64, u32, vp]
    L.og_t_mover_threshold.restype = u32
    L.og_t_mover_threshold.argtypes = [i64, i64]
    L.og_t_below_0999.restype = ci
    L.og_t_below_0999.argtypes = [u64, u64]
    L.og_t_proposals.argtypes = [i64, vp, vp, vp, vp, vp, vp, u64, u32, ci, ci, ci, u32, ci, vp, vp, vp, vp, vp]
    L.og_t_pack_isolated.argtypes = [i64, vp, vp, u64, vp, vp]
    L.og_t_two_hop.argtypes = [i64, vp, vp, vp, vp, vp, vp, u64, u32, ci, i64]
    L.og_t_lp_propose.restype = i64
    L.og_t_lp_propose.argtypes = [i64, vp, vp, vp, vp, ci, vp, vp, vp, u64, vp, dbl, vp, vp]
    L.og_t_afterburner.restype = i64
    L.og_t_afterburner.argtypes = [i64, vp, vp, vp, vp, ci, vp, vp, u64, vp, vp, vp, vp]
    L.og_t_gain_bucket.argtypes = [dbl]
    L.og_t_repair_destination.argtypes = [i64, vp, vp, vp, vp, ci, vp, vp, u64, ctypes.c_int32, vp, vp]
    L.og_t_repair_iteration.restype = i64
    L.og_t_repair_iteration.argtypes = [i64, vp, vp, vp, vp, ci, vp, vp, u64, vp, vp, vp, vp]
    _STAGES_BOUND = True


def _graph_arrays(off, nbr, ew, vw):
    off = np.ascontiguousarray(off, np.int32)
    nbr = np.ascontiguousarray(nbr if nbr is not None else np.zeros(0, np.int32), np.int32)
    ew = None if ew is None else np.ascontiguousarray(ew, np.uint32)
    n = len(off) - 1
    vw = np.ones(n, np.uint32) if vw is None else np.ascontiguousarray(vw, np.uint32)
    return n, off, nbr, ew, vw


def seed_match(off, nbr, ew, vw, cap, salt):
    """Seed matching: (cluster root per vertex, cluster count)."""
    _bind_stages()
    n, off, nbr, ew, vw = _graph_arrays(off, nbr, ew, vw)
    roots = np.zeros(n, np.int32)
    count = lib().og_t_seed_match(n, _p(off), _p(nbr), _p(ew), _p(vw), cap, salt, _p(roots))
    return roots, int(count)


def below_0999(cut, best):
    _bind_stages()
    return bool(lib().og_t_below_0999(cut, best))


def mover_threshold(clusters, n):
    _bind_stages()
    return int(lib().og_t_mover_threshold(clusters, n))


def proposals(off, nbr, ew, vw, clusters, cw, cap, seed, level, round_, filter_roles, threshold, need_regret):
    """Roles and top-two proposals: dict of roles, proposals, gains, regret, ties."""
    _bind_stages()
    n, off, nbr, ew, vw = _graph_arrays(off, nbr, ew, vw)
    clusters = np.ascontiguousarray(clusters, np.int32)
    cw = np.ascontiguousarray(cw, np.uint64)
    out = dict(roles=np.zeros(n, np.uint8), proposals=np.zeros(n, np.int32), gains=np.zeros(n, np.uint64),
               regret=np.zeros(n, np.uint64), ties=np.zeros(n, np.uint32))
    lib().og_t_proposals(n, _p(off), _p(nbr), _p(ew), _p(vw), _p(clusters), _p(cw), cap, seed, level, round_,
                         filter_roles, threshold, need_regret, _p(out["roles"]), _p(out["proposals"]),
                         _p(out["gains"]), _p(out["regret"]), _p(out["ties"]))
    return out


def pack_isolated(off, vw, cap, clusters, cw):
    """Isolated packing in place on copies: (clusters, cw, ran)."""
    _bind_stages()
    n, off, _, _, vw = _graph_arrays(off, None, None, vw)
    clusters = np.array(clusters, np.int32, copy=True)
    cw = np.array(cw, np.uint64, copy=True)
    ran = lib().og_t_pack_isolated(n, _p(off), _p(vw), cap, _p(clusters), _p(cw))
    return clusters, cw, bool(ran)


def two_hop(off, nbr, ew, vw, clusters, cw, cap, seed, level, budget):
    _bind_stages()
    n, off, nbr, ew, vw = _graph_arrays(off, nbr, ew, vw)
    clusters = np.array(clusters, np.int32, copy=True)
    cw = np.array(cw, np.uint64, copy=True)
    lib().og_t_two_hop(n, _p(off), _p(nbr), _p(ew), _p(vw), _p(clusters), _p(cw), cap, seed, level, budget)
    return clusters, cw


def lp_propose(off, nbr, ew, vw, k, labels, locks, loads, cap, prices, filter_):
    """One LP proposal sweep: (targets, gains, locks after, candidate count)."""
    _bind_stages()
    n, off, nbr, ew, vw = _graph_arrays(off, nbr, ew, vw)
    labels = np.ascontiguousarray(labels, np.int32)
    locks = np.array(locks, np.int32, copy=True)
    loads = np.ascontiguousarray(loads, np.uint64)
    prices = None if prices is None else np.ascontiguousarray(prices, np.float64)
    targets = np.zeros(n, np.int32)
    gains = np.zeros(n, np.int64)
    c = lib().og_t_lp_propose(n, _p(off), _p(nbr), _p(ew), _p(vw), k, _p(labels), _p(locks), _p(loads), cap,
                              _p(prices), filter_, _p(targets), _p(gains))
    return targets, gains, locks, int(c)


def afterburner(off, nbr, ew, vw, k, labels, loads, cap, prices, targets, gains):
    """Afterburner over candidates (targets >= 0): keep flags."""
    _bind_stages()
    n, off, nbr, ew, vw = _graph_arrays(off, nbr, ew, vw)
    labels = np.ascontiguousarray(labels, np.int32)
    loads = np.ascontiguousarray(loads, np.uint64)
    prices = None if prices is None else np.ascontiguousarray(prices, np.float64)
    targets = np.ascontiguousarray(targets, np.int32)
    gains = np.ascontiguousarray(gains, np.int64)
    keep = np.zeros(n, np.uint8)
    lib().og_t_afterburner(n, _p(off), _p(nbr), _p(ew), _p(vw), k, _p(labels), _p(loads), cap, _p(prices),
                           _p(targets), _p(gains), _p(keep))
    return keep.astype(bool)


def gain_bucket(gain):
    _bind_stages()
    return int(lib().og_t_gain_bucket(float(gain)))


def repair_destination(off, nbr, ew, vw, k, labels, loads, cap, v):
    """(destination or -1, A(v, destination), A(v, own part))."""
    _bind_stages()
    n, off, nbr, ew, vw = _graph_arrays(off, nbr, ew, vw)
    labels = np.ascontiguousarray(labels, np.int32)
    loads = np.ascontiguousarray(loads, np.uint64)
    b = ctypes.c_uint64(); c = ctypes.c_uint64()
    d = lib().og_t_repair_destination(n, _p(off), _p(nbr), _p(ew), _p(vw), k, _p(labels), _p(loads), cap, v,
                                      ctypes.byref(b), ctypes.byref(c))
    return int(d), b.value, c.value


def repair_iteration(off, nbr, ew, vw, k, labels, loads, cap):
    """One repair iteration, not applied: dict of first_choice, bucket, eligible, receiver."""
    _bind_stages()
    n, off, nbr, ew, vw = _graph_arrays(off, nbr, ew, vw)
    labels = np.ascontiguousarray(labels, np.int32)
    loads = np.ascontiguousarray(loads, np.uint64)
    out = dict(first_choice=np.zeros(n, np.int32), bucket=np.zeros(n, np.int32), eligible=np.zeros(n, np.uint8),
               receiver=np.zeros(n, np.int32))
    lib().og_t_repair_iteration(n, _p(off), _p(nbr), _p(ew), _p(vw), k, _p(labels), _p(loads), cap,
                                _p(out["first_choice"]), _p(out["bucket"]), _p(out["eligible"]),
                                _p(out["receiver"]))
    out["eligible"] = out["eligible"].astype(bool)
    return out
