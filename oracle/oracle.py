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
