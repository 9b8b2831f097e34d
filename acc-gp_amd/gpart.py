"""ctypes binding of libgpart_mi355x.so (include/gpart_mi355x.h).

Mirrors the reference's C++ entry points (coarsen, initial_partition_metis,
uncoarsen, refine_partition, partition_capacity) for Python callers.  Raises
RuntimeError if the library has not been built; there is no fallback path.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgpart_mi355x.so")
_LIB = None

GP_OK = 0


class CoarsenOptions(ctypes.Structure):
    _fields_ = [
        ("parts", ctypes.c_int), ("seed", ctypes.c_uint32), ("stop_ratio", ctypes.c_double),
        ("max_levels", ctypes.c_int), ("cluster_size_factor", ctypes.c_int), ("pack_isolated", ctypes.c_int),
        ("progressive", ctypes.c_int), ("fine_factor", ctypes.c_int), ("capacity_growth", ctypes.c_int),
        ("seed_matching", ctypes.c_int), ("fast_path", ctypes.c_int), ("hem_degree_limit", ctypes.c_int),
    ]


class RefineOptions(ctypes.Structure):
    _fields_ = [
        ("parts", ctypes.c_int), ("imbalance", ctypes.c_double), ("max_rounds", ctypes.c_int),
        ("stale_rounds", ctypes.c_int), ("growth_ratio", ctypes.c_double), ("exchange_rounds", ctypes.c_int),
        ("strict_verify", ctypes.c_int), ("gpu_lp_loop", ctypes.c_int),
    ]


class LevelStats(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64), ("m", ctypes.c_int64), ("refined", ctypes.c_int),
        ("initial_repair_failed", ctypes.c_int), ("rounds", ctypes.c_int),
        ("initial_cut", ctypes.c_uint64), ("final_cut", ctypes.c_uint64),
        ("lp_moves", ctypes.c_uint64), ("repair_moves", ctypes.c_uint64), ("group_moves", ctypes.c_uint64),
        ("repair_iterations", ctypes.c_int), ("strict_checks", ctypes.c_uint64),
    ]


class PartitionStats(ctypes.Structure):
    _fields_ = [
        ("coarsen_seconds", ctypes.c_double), ("initial_partition_seconds", ctypes.c_double),
        ("uncoarsen_seconds", ctypes.c_double), ("algorithm_total_seconds", ctypes.c_double),
        ("levels", ctypes.c_int), ("coarsest_vertices", ctypes.c_int64), ("final_cut", ctypes.c_uint64),
        ("max_part_weight", ctypes.c_uint64), ("capacity", ctypes.c_uint64),
        ("total_edge_weight", ctypes.c_uint64), ("launches", ctypes.c_uint64), ("host_syncs", ctypes.c_uint64),
    ]


class Request(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64), ("m", ctypes.c_int64), ("offsets", ctypes.c_void_p), ("neighbors", ctypes.c_void_p),
        ("edge_weights", ctypes.c_void_p), ("vertex_weights", ctypes.c_void_p), ("parts", ctypes.c_int),
        ("imbalance", ctypes.c_double), ("seed", ctypes.c_uint32), ("warm_labels", ctypes.c_void_p),
        ("labels", ctypes.c_void_p), ("graph_id", ctypes.c_uint64), ("delta_count", ctypes.c_int64),
        ("delta_u", ctypes.c_void_p), ("delta_v", ctypes.c_void_p), ("delta_w", ctypes.c_void_p),
        ("delta_insert", ctypes.c_void_p), ("warm_resident", ctypes.c_int),
        ("stats", PartitionStats), ("status", ctypes.c_int),
        ("queue_seconds", ctypes.c_double), ("device_seconds", ctypes.c_double), ("error", ctypes.c_char * 256),
    ]


# Every symbol declared in include/gpart_mi355x.h.
EXPORTS = [
    "gp_last_error", "gp_context_create", "gp_context_destroy", "gp_context_reserved_bytes", "gp_context_set_overlap",
    "gp_graph_create", "gp_graph_info", "gp_graph_download", "gp_graph_metadata", "gp_graph_destroy", "gp_capacity",
    "gp_default_coarsen_options", "gp_coarsen", "gp_hierarchy_levels", "gp_hierarchy_level",
    "gp_hierarchy_map_download", "gp_hierarchy_level_info", "gp_hierarchy_assemble", "gp_graph_edge_weight",
    "gp_hierarchy_level_capacity", "gp_hierarchy_stop_reason",
    "gp_hierarchy_destroy", "gp_workspace_create", "gp_workspace_destroy", "gp_aggregate_level", "gp_aggregate_info",
    "gp_aggregate_map_download", "gp_aggregate_vertex_weights_download", "gp_aggregate_destroy", "gp_contract",
    "gp_initial_partition_host", "gp_initial_partition", "gp_repair_capacity_host", "gp_initial_partition_backend",
    "gp_default_refine_options", "gp_uncoarsen",
    "gp_result_labels", "gp_result_cut", "gp_result_loads", "gp_result_levels", "gp_result_level_stats",
    "gp_result_destroy", "gp_result_device_labels", "gp_edge_cut", "gp_edge_cut_noalloc", "gp_refine_partition", "gp_partition", "gp_generate_grid", "gp_generate_rmat",
    "gp_generate_rgg2d", "gp_generate_knn3d", "gp_free_host", "gp_load_csr", "gp_save_csr",
    "gp_profile_enable", "gp_profile_query", "gp_profile_reset", "gp_server_create", "gp_server_submit",
    "gp_server_wait", "gp_server_destroy", "gp_partition_batch", "gp_server_graph_add", "gp_server_graph_download",
    "gp_server_graph_release", "gp_graph_apply_delta", "gp_debug_below_0999",
    "gp_debug_radix_sort",
]


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libgpart_mi355x.so is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, u64, i32, u32, dbl = (ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64, ctypes.c_int,
                                       ctypes.c_uint32, ctypes.c_double)
        pp = ctypes.POINTER(ctypes.c_void_p)
        L.gp_last_error.restype = ctypes.c_char_p
        L.gp_context_create.argtypes = [i32, pp]
        L.gp_context_destroy.argtypes = [vp]
        L.gp_context_reserved_bytes.argtypes = [vp]
        L.gp_context_reserved_bytes.restype = u64
        L.gp_context_set_overlap.argtypes = [vp, i32]
        L.gp_graph_create.argtypes = [vp, i64, i64, vp, vp, vp, vp, pp]
        L.gp_graph_info.argtypes = [vp, vp, vp]
        L.gp_graph_download.argtypes = [vp, vp, vp, vp, vp, vp]
        L.gp_graph_destroy.argtypes = [vp]
        L.gp_graph_metadata.argtypes = [vp, vp, vp, vp, vp]
        L.gp_capacity.argtypes = [u64, i32, dbl, vp]
        L.gp_default_coarsen_options.argtypes = [i32, u32, vp]
        L.gp_coarsen.argtypes = [vp, vp, vp, pp]
        L.gp_hierarchy_levels.argtypes = [vp]
        L.gp_hierarchy_level.argtypes = [vp, i32]
        L.gp_hierarchy_level.restype = vp
        L.gp_hierarchy_map_download.argtypes = [vp, vp, i32, vp]
        L.gp_hierarchy_level_info.argtypes = [vp, i32, vp, vp, vp]
        L.gp_hierarchy_level_capacity.argtypes = [vp, i32, vp]
        L.gp_hierarchy_assemble.argtypes = [vp, i32, vp, vp, pp]
        L.gp_graph_edge_weight.argtypes = [vp, vp, vp]
        L.gp_hierarchy_stop_reason.argtypes = [vp]
        L.gp_hierarchy_stop_reason.restype = ctypes.c_char_p
        L.gp_hierarchy_destroy.argtypes = [vp]
        L.gp_workspace_create.argtypes = [pp]
        L.gp_workspace_destroy.argtypes = [vp]
        L.gp_aggregate_level.argtypes = [vp, vp, vp, i32, u32, i32, i32, i32, i32, i32, pp]
        L.gp_aggregate_info.argtypes = [vp, vp, vp, vp]
        L.gp_aggregate_map_download.argtypes = [vp, vp, vp]
        L.gp_aggregate_vertex_weights_download.argtypes = [vp, vp, vp]
        L.gp_aggregate_destroy.argtypes = [vp]
        L.gp_contract.argtypes = [vp, vp, vp, vp, pp]
        L.gp_initial_partition_host.argtypes = [i64, vp, vp, vp, vp, i32, dbl, u32, i32, vp]
        L.gp_initial_partition.argtypes = [vp, vp, i32, dbl, u32, i32, vp]
        L.gp_repair_capacity_host.argtypes = [i64, vp, vp, vp, vp, i32, u64, vp, vp]
        L.gp_initial_partition_backend.restype = ctypes.c_char_p
        L.gp_default_refine_options.argtypes = [i32, dbl, vp]
        L.gp_uncoarsen.argtypes = [vp, vp, vp, vp, pp]
        L.gp_result_labels.argtypes = [vp, vp, vp]
        L.gp_result_cut.argtypes = [vp]
        L.gp_result_cut.restype = u64
        L.gp_result_loads.argtypes = [vp, vp]
        L.gp_result_levels.argtypes = [vp]
        L.gp_result_level_stats.argtypes = [vp, i32, vp]
        L.gp_result_destroy.argtypes = [vp]
        L.gp_result_device_labels.argtypes = [vp]
        L.gp_result_device_labels.restype = vp
        L.gp_edge_cut.argtypes = [vp, vp, vp, i32, vp, vp]
        L.gp_edge_cut_noalloc.argtypes = [vp, vp, vp, i32, vp, vp]
        L.gp_refine_partition.argtypes = [vp, vp, vp, vp, vp]
        L.gp_partition.argtypes = [vp, vp, i32, dbl, u32, vp, vp]
        L.gp_generate_grid.argtypes = [i64, i64, vp, vp, vp, vp]
        L.gp_generate_rmat.argtypes = [i32, i32, dbl, dbl, dbl, u64, vp, vp, vp, vp]
        L.gp_generate_rgg2d.argtypes = [i64, dbl, u64, vp, vp, vp, vp]
        L.gp_generate_knn3d.argtypes = [i64, i32, u64, vp, vp, vp, vp, vp, vp]
        L.gp_free_host.argtypes = [vp]
        L.gp_load_csr.argtypes = [ctypes.c_char_p, ctypes.c_char_p, vp, vp, vp, vp]
        L.gp_save_csr.argtypes = [ctypes.c_char_p, ctypes.c_char_p, i64, i64, vp, vp]
        L.gp_profile_enable.argtypes = [vp, i32]
        L.gp_profile_query.argtypes = [vp, ctypes.c_char_p, vp, vp, vp]
        L.gp_profile_reset.argtypes = [vp]
        L.gp_server_create.argtypes = [i32, i32, u64, pp]
        L.gp_server_submit.argtypes = [vp, vp, pp]
        L.gp_server_wait.argtypes = [vp, vp]
        L.gp_server_destroy.argtypes = [vp]
        L.gp_partition_batch.argtypes = [i32, vp, i64, i32, u64]
        L.gp_server_graph_add.argtypes = [vp, i64, i64, vp, vp, vp, vp, vp]
        L.gp_server_graph_download.argtypes = [vp, u64, vp, vp, vp, vp, vp, vp]
        L.gp_server_graph_release.argtypes = [vp, u64]
        L.gp_graph_apply_delta.argtypes = [vp, vp, i64, vp, vp, vp, vp, vp]
        L.gp_debug_below_0999.argtypes = [u64, u64]
        L.gp_debug_radix_sort.argtypes = [vp, vp, vp, i64, i32, i32]
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def check(rc):
    if rc != GP_OK:
        raise RuntimeError("gpart error %d: %s" % (rc, lib().gp_last_error().decode()))


def below_0999(cut, best):
    """The device loop control's integer form of the small-level stale rule
    cut < 0.999 * best (long double)."""
    return bool(lib().gp_debug_below_0999(cut, best))


def radix_sort(ctx, keys, vals=None, begin_bit=0, end_bit=64):
    """The device radix sort behind the admission and contraction orders:
    stable by bits [begin_bit, end_bit) of the uint64 keys, uint32 vals
    carried along.  The passes take whole 8-bit digits from begin_bit, so
    key bits past end_bit in the last digit must be zero.  Returns sorted
    copies."""
    k = np.ascontiguousarray(keys, dtype=np.uint64).copy()
    v = None if vals is None else np.ascontiguousarray(vals, dtype=np.uint32).copy()
    check(lib().gp_debug_radix_sort(ctx.h, _p(k), _p(v), len(k), begin_bit, end_bit))
    return k, v


def capacity(total, parts, imbalance=1.03):
    out = ctypes.c_uint64()
    check(lib().gp_capacity(total, parts, imbalance, ctypes.byref(out)))
    return out.value


# ------------------------------------------------------------- host graphs
def _take_csr(n, m, off_p, nbr_p):
    off = np.ctypeslib.as_array(ctypes.cast(off_p, ctypes.POINTER(ctypes.c_int32)), shape=(n.value + 1,)).copy()
    nbr = np.ctypeslib.as_array(ctypes.cast(nbr_p, ctypes.POINTER(ctypes.c_int32)),
                                shape=(max(m.value, 1),))[:m.value].copy()
    lib().gp_free_host(off_p)
    lib().gp_free_host(nbr_p)
    return off, nbr


def generate_grid(rows, cols):
    n, m, a, b = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_void_p(), ctypes.c_void_p()
    check(lib().gp_generate_grid(rows, cols, ctypes.byref(n), ctypes.byref(m), ctypes.byref(a), ctypes.byref(b)))
    return _take_csr(n, m, a, b)


def generate_rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1):
    n, m, x, y = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_void_p(), ctypes.c_void_p()
    check(lib().gp_generate_rmat(scale, edge_factor, a, b, c, seed, ctypes.byref(n), ctypes.byref(m),
                                 ctypes.byref(x), ctypes.byref(y)))
    return _take_csr(n, m, x, y)


def generate_rgg2d(points, radius, seed=5):
    n, m, x, y = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_void_p(), ctypes.c_void_p()
    check(lib().gp_generate_rgg2d(points, radius, seed, ctypes.byref(n), ctypes.byref(m), ctypes.byref(x),
                                  ctypes.byref(y)))
    return _take_csr(n, m, x, y)


def generate_knn3d(points, knn=10, seed=2):
    n, m, x, y, ew, vw = (ctypes.c_int64(), ctypes.c_int64(), ctypes.c_void_p(), ctypes.c_void_p(),
                          ctypes.c_void_p(), ctypes.c_void_p())
    check(lib().gp_generate_knn3d(points, knn, seed, ctypes.byref(n), ctypes.byref(m), ctypes.byref(x),
                                  ctypes.byref(y), ctypes.byref(ew), ctypes.byref(vw)))
    mm = m.value
    e = np.ctypeslib.as_array(ctypes.cast(ew, ctypes.POINTER(ctypes.c_uint32)), shape=(max(mm, 1),))[:mm].copy()
    v = np.ctypeslib.as_array(ctypes.cast(vw, ctypes.POINTER(ctypes.c_uint32)), shape=(n.value,)).copy()
    lib().gp_free_host(ew)
    lib().gp_free_host(vw)
    off, nbr = _take_csr(n, m, x, y)
    return off, nbr, e, v


def load_csr(indptr, indices):
    n, m, x, y = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_void_p(), ctypes.c_void_p()
    check(lib().gp_load_csr(indptr.encode(), indices.encode(), ctypes.byref(n), ctypes.byref(m),
                            ctypes.byref(x), ctypes.byref(y)))
    return _take_csr(n, m, x, y)


def save_csr(indptr, indices, off, nbr):
    off = np.ascontiguousarray(off, np.int32)
    nbr = np.ascontiguousarray(nbr, np.int32)
    check(lib().gp_save_csr(indptr.encode(), indices.encode(), len(off) - 1, len(nbr), _p(off), _p(nbr)))


def initial_partition_host(off, nbr, ew, vw, parts, imbalance=1.03, seed=0, require_feasible=False):
    """initial_partition_metis on a host CSR; require_feasible=False is the
    production setting (app/partition_main.cu:111-112)."""
    off = np.ascontiguousarray(off, np.int32)
    nbr = np.ascontiguousarray(nbr, np.int32)
    ew = None if ew is None else np.ascontiguousarray(ew, np.uint32)
    vw = None if vw is None else np.ascontiguousarray(vw, np.uint32)
    out = np.zeros(len(off) - 1, np.int32)
    check(lib().gp_initial_partition_host(len(off) - 1, _p(off), _p(nbr), _p(ew), _p(vw), parts, imbalance, seed,
                                          int(require_feasible), _p(out)))
    return out


def repair_capacity_host(off, nbr, ew, vw, parts, capacity, labels):
    """Strict single-move capacity repair: (labels after, vertices moved)."""
    off = np.ascontiguousarray(off, np.int32)
    nbr = np.ascontiguousarray(nbr, np.int32)
    ew = None if ew is None else np.ascontiguousarray(ew, np.uint32)
    vw = None if vw is None else np.ascontiguousarray(vw, np.uint32)
    out = np.array(labels, np.int32, copy=True)
    moved = ctypes.c_uint64()
    check(lib().gp_repair_capacity_host(len(off) - 1, _p(off), _p(nbr), _p(ew), _p(vw), parts, capacity, _p(out),
                                        ctypes.byref(moved)))
    return out, moved.value


def initial_partition_backend():
    return lib().gp_initial_partition_backend().decode()


def coarsen_options(parts, seed=0, **overrides):
    o = CoarsenOptions()
    lib().gp_default_coarsen_options(parts, seed, ctypes.byref(o))
    for key, value in overrides.items():
        setattr(o, key, value)
    return o


def refine_options(parts, imbalance=1.03, **overrides):
    o = RefineOptions()
    lib().gp_default_refine_options(parts, imbalance, ctypes.byref(o))
    for key, value in overrides.items():
        setattr(o, key, value)
    return o


# ------------------------------------------------------------ device side
class Context:
    def __init__(self, device=0):
        self.h = ctypes.c_void_p()
        check(lib().gp_context_create(device, ctypes.byref(self.h)))

    def close(self):
        if self.h:
            lib().gp_context_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def profile(self, enable=True):
        check(lib().gp_profile_enable(self.h, int(enable)))

    def set_overlap(self, on=True):
        check(lib().gp_context_set_overlap(self.h, int(on)))

    def profile_query(self, name):
        ms, by, ln = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        check(lib().gp_profile_query(self.h, name.encode(), ctypes.byref(ms), ctypes.byref(by), ctypes.byref(ln)))
        return ms.value, by.value, ln.value

    def profile_reset(self):
        check(lib().gp_profile_reset(self.h))


class Graph:
    def __init__(self, ctx, off, nbr, ew=None, vw=None):
        self.ctx = ctx
        off = np.ascontiguousarray(off, np.int32)
        nbr = np.ascontiguousarray(nbr, np.int32)
        ew = None if ew is None else np.ascontiguousarray(ew, np.uint32)
        vw = None if vw is None else np.ascontiguousarray(vw, np.uint32)
        self.n = len(off) - 1
        self.m = len(nbr)
        self.h = ctypes.c_void_p()
        check(lib().gp_graph_create(ctx.h, self.n, self.m, _p(off), _p(nbr), _p(ew), _p(vw), ctypes.byref(self.h)))

    def close(self):
        if self.h:
            lib().gp_graph_destroy(self.h)
            self.h = ctypes.c_void_p()

    def partition(self, parts, imbalance=1.03, seed=0, download=True):
        labels = np.zeros(self.n, np.int32) if download else None
        st = PartitionStats()
        check(lib().gp_partition(self.ctx.h, self.h, parts, imbalance, seed, _p(labels), ctypes.byref(st)))
        return labels, dict((f, getattr(st, f)) for f, _ in PartitionStats._fields_)

    def edge_cut(self, device_labels, parts):
        """(cut, loads) of labels already in device memory (gp_edge_cut)."""
        return _edge_cut(self.ctx, self.h, device_labels, parts)

    def refine(self, labels, ropts):
        out = np.array(labels, np.int32, copy=True)
        st = LevelStats()
        check(lib().gp_refine_partition(self.ctx.h, self.h, _p(out), ctypes.byref(ropts), ctypes.byref(st)))
        return out, dict((f, getattr(st, f)) for f, _ in LevelStats._fields_)


def edge_cut_noalloc(ctx, g_handle, device_labels, parts):
    """(cut, loads) of device labels through gp_edge_cut_noalloc on ctx."""
    cut = ctypes.c_uint64()
    loads = np.zeros(parts, np.uint64)
    check(lib().gp_edge_cut_noalloc(ctx.h, g_handle, ctypes.c_void_p(device_labels), parts, ctypes.byref(cut),
                                    _p(loads)))
    return cut.value, loads


def edge_cut(ctx, g_handle, device_labels, parts):
    """(cut, loads) of device labels through gp_edge_cut on ctx."""
    return _edge_cut(ctx, g_handle, device_labels, parts)


def reserved_bytes(ctx):
    return lib().gp_context_reserved_bytes(ctx.h)


def _edge_cut(ctx, g_handle, device_labels, parts):
    cut = ctypes.c_uint64()
    loads = np.zeros(parts, np.uint64)
    check(lib().gp_edge_cut(ctx.h, g_handle, ctypes.c_void_p(device_labels), parts, ctypes.byref(cut), _p(loads)))
    return cut.value, loads


def download_graph(ctx, g_handle):
    n, m = ctypes.c_int64(), ctypes.c_int64()
    check(lib().gp_graph_info(g_handle, ctypes.byref(n), ctypes.byref(m)))
    off = np.zeros(n.value + 1, np.int32)
    nbr = np.zeros(max(m.value, 1), np.int32)
    ew = np.zeros(max(m.value, 1), np.uint32)
    vw = np.zeros(n.value, np.uint32)
    check(lib().gp_graph_download(ctx.h, g_handle, _p(off), _p(nbr), _p(ew), _p(vw)))
    return off, nbr[:m.value], ew[:m.value], vw


class Hierarchy:
    def __init__(self, ctx, graph, opts):
        self.ctx = ctx
        self.graph = graph
        self.h = ctypes.c_void_p()
        check(lib().gp_coarsen(ctx.h, graph.h, ctypes.byref(opts), ctypes.byref(self.h)))

    def close(self):
        if self.h:
            lib().gp_hierarchy_destroy(self.h)
            self.h = ctypes.c_void_p()

    @property
    def levels(self):
        return lib().gp_hierarchy_levels(self.h)

    def graph_at(self, level):
        return download_graph(self.ctx, lib().gp_hierarchy_level(self.h, level))

    def map(self, level):
        n = ctypes.c_int64()
        check(lib().gp_graph_info(lib().gp_hierarchy_level(self.h, level), ctypes.byref(n), None))
        out = np.zeros(n.value, np.int32)
        check(lib().gp_hierarchy_map_download(self.ctx.h, self.h, level, _p(out)))
        return out

    def info(self, level):
        md, hh, ib = ctypes.c_uint64(), ctypes.c_int(), ctypes.c_uint64()
        check(lib().gp_hierarchy_level_info(self.h, level, ctypes.byref(md), ctypes.byref(hh), ctypes.byref(ib)))
        return dict(max_degree=md.value, has_heavy=hh.value, incident_bound=ib.value)

    def uncoarsen(self, coarsest_labels, ropts, evaluate_parts=None, evaluate=None):
        """Refined labels, cut, loads and per-level stats.  With
        evaluate_parts, also the (cut, loads) that gp_edge_cut counts on the
        result's device labels with that many parts (appended); with
        evaluate, what evaluate(graph_handle, device_labels) returns while
        the result is alive (appended)."""
        labels = np.ascontiguousarray(coarsest_labels, np.int32)
        r = ctypes.c_void_p()
        check(lib().gp_uncoarsen(self.ctx.h, self.h, _p(labels), ctypes.byref(ropts), ctypes.byref(r)))
        try:
            evaluated = None
            if evaluate is not None:
                evaluated = evaluate(self.graph.h, lib().gp_result_device_labels(r))
            elif evaluate_parts is not None:
                evaluated = _edge_cut(self.ctx, self.graph.h, lib().gp_result_device_labels(r), evaluate_parts)
            out = np.zeros(self.graph.n, np.int32)
            check(lib().gp_result_labels(self.ctx.h, r, _p(out)))
            loads = np.zeros(ropts.parts, np.uint64)
            check(lib().gp_result_loads(r, _p(loads)))
            stats = []
            for i in range(lib().gp_result_levels(r)):
                st = LevelStats()
                check(lib().gp_result_level_stats(r, i, ctypes.byref(st)))
                stats.append(dict((f, getattr(st, f)) for f, _ in LevelStats._fields_))
            if evaluate is not None or evaluate_parts is not None:
                return out, lib().gp_result_cut(r), loads, stats, evaluated
            return out, lib().gp_result_cut(r), loads, stats
        finally:
            lib().gp_result_destroy(r)


class BatchItem:
    """One request of a batch or of the server: a host CSR and its options."""

    def __init__(self, off, nbr, parts, imbalance=1.03, seed=0, ew=None, vw=None, warm=None):
        self.off = np.ascontiguousarray(off, np.int32)
        self.nbr = np.ascontiguousarray(nbr, np.int32)
        self.ew = None if ew is None else np.ascontiguousarray(ew, np.uint32)
        self.vw = None if vw is None else np.ascontiguousarray(vw, np.uint32)
        self.warm = None if warm is None else np.ascontiguousarray(warm, np.int32)
        self.labels = np.zeros(len(self.off) - 1, np.int32)
        self.parts, self.imbalance, self.seed = parts, imbalance, seed

    def fill(self, req):
        req.n, req.m = len(self.off) - 1, len(self.nbr)
        req.offsets, req.neighbors = _p(self.off), _p(self.nbr)
        req.edge_weights, req.vertex_weights = _p(self.ew), _p(self.vw)
        req.parts, req.imbalance, req.seed = self.parts, self.imbalance, self.seed
        req.warm_labels, req.labels = _p(self.warm), _p(self.labels)


class ResidentItem:
    """A request on a server's resident graph: optional edge changes
    (arrays u, v, insert flags, weights), then a cold partition at `parts`,
    or with warm=True a refinement from the graph's current partition."""

    def __init__(self, graph_id, n, parts, imbalance=1.03, seed=0, delta=None, warm=False):
        self.graph_id, self.parts, self.imbalance, self.seed, self.warm = graph_id, parts, imbalance, seed, warm
        self.labels = np.zeros(n, np.int32)
        self.delta = None
        if delta is not None:
            u, v, ins, w = delta
            self.delta = (np.ascontiguousarray(u, np.int32), np.ascontiguousarray(v, np.int32),
                          np.ascontiguousarray(ins, np.uint8), None if w is None else np.ascontiguousarray(w, np.uint32))

    def fill(self, req):
        req.graph_id = self.graph_id
        req.parts, req.imbalance, req.seed = self.parts, self.imbalance, self.seed
        req.labels = _p(self.labels)
        req.warm_resident = int(self.warm)
        if self.delta is not None:
            u, v, ins, w = self.delta
            req.delta_count = len(u)
            req.delta_u, req.delta_v, req.delta_insert, req.delta_w = _p(u), _p(v), _p(ins), _p(w)


def apply_delta(ctx, g_handle, u, v, insert, w=None):
    """gp_graph_apply_delta on a device graph; returns the new entry count."""
    u = np.ascontiguousarray(u, np.int32)
    v = np.ascontiguousarray(v, np.int32)
    insert = np.ascontiguousarray(insert, np.uint8)
    w = None if w is None else np.ascontiguousarray(w, np.uint32)
    m = ctypes.c_int64()
    check(lib().gp_graph_apply_delta(ctx.h, g_handle, len(u), _p(u), _p(v), _p(w), _p(insert), ctypes.byref(m)))
    return m.value


def _stats_dict(req):
    return dict((f, getattr(req.stats, f)) for f, _ in PartitionStats._fields_)


def partition_batch(items, workers=6, hbm_budget=0, device=0):
    """Partitions BatchItems concurrently; returns per-item stats dicts."""
    reqs = (Request * len(items))()
    for it, r in zip(items, reqs):
        it.fill(r)
    rc = lib().gp_partition_batch(device, ctypes.cast(reqs, ctypes.c_void_p), len(items), workers, hbm_budget)
    for r in reqs:
        if r.status != GP_OK:
            raise RuntimeError("gpart batch item error %d: %s" % (r.status, r.error.decode()))
    check(rc)
    return [dict(_stats_dict(r), queue_seconds=r.queue_seconds, device_seconds=r.device_seconds) for r in reqs]


class Server:
    """Continuous-batching request loop (gp_server_*)."""

    def __init__(self, workers=6, hbm_budget=0, device=0):
        self.h = ctypes.c_void_p()
        check(lib().gp_server_create(device, workers, hbm_budget, ctypes.byref(self.h)))
        self._live = {}

    def submit(self, item):
        req = Request()
        item.fill(req)
        t = ctypes.c_void_p()
        check(lib().gp_server_submit(self.h, ctypes.byref(req), ctypes.byref(t)))
        self._live[t.value] = (req, item)
        return t.value

    def add_graph(self, off, nbr, ew=None, vw=None):
        """Keeps a graph resident in HBM; returns its id."""
        off = np.ascontiguousarray(off, np.int32)
        nbr = np.ascontiguousarray(nbr, np.int32)
        ew = None if ew is None else np.ascontiguousarray(ew, np.uint32)
        vw = None if vw is None else np.ascontiguousarray(vw, np.uint32)
        gid = ctypes.c_uint64()
        check(lib().gp_server_graph_add(self.h, len(off) - 1, len(nbr), _p(off), _p(nbr), _p(ew), _p(vw),
                                        ctypes.byref(gid)))
        return gid.value

    def graph(self, graph_id):
        """(offsets, neighbors, edge weights, current labels) of a resident graph."""
        n, m = ctypes.c_int64(), ctypes.c_int64()
        check(lib().gp_server_graph_download(self.h, graph_id, ctypes.byref(n), ctypes.byref(m), None, None, None,
                                             None))
        off = np.zeros(n.value + 1, np.int32)
        nbr = np.zeros(max(m.value, 1), np.int32)
        ew = np.zeros(max(m.value, 1), np.uint32)
        lab = np.full(n.value, -1, np.int32)
        check(lib().gp_server_graph_download(self.h, graph_id, None, None, _p(off), _p(nbr), _p(ew), _p(lab)))
        return off, nbr[:m.value], ew[:m.value], lab

    def release_graph(self, graph_id):
        check(lib().gp_server_graph_release(self.h, graph_id))

    def wait(self, ticket):
        req, item = self._live.pop(ticket)
        rc = lib().gp_server_wait(self.h, ctypes.c_void_p(ticket))
        if rc != GP_OK:
            raise RuntimeError("gpart request error %d: %s" % (rc, req.error.decode()))
        return dict(_stats_dict(req), queue_seconds=req.queue_seconds, device_seconds=req.device_seconds)

    def close(self):
        """Finishes every queued request, stops the workers and releases
        tickets not waited for.  The request buffers stay referenced until
        the workers have stopped."""
        if self.h:
            lib().gp_server_destroy(self.h)
            self.h = None
        self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        if getattr(self, "h", None):
            self.close()
