"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Deterministic mode is the only mode, so every comparison is bit-exact:
coarsening maps and coarse CSR arrays per level, refined labels, cut and
part loads.  Full-size configurations are covered by size-independent
properties in test_gpu_properties.py.
"""
import numpy as np
import pytest

from tests import graphs

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _cases(gp):
    cases = []
    off, nbr = graphs.grid(32, 32)
    cases.append(("grid32_k4", off, nbr, None, None, 4))
    off, nbr = gp.generate_grid(256, 256)
    cases.append(("grid256_k8", off, nbr, None, None, 8))
    off, nbr = gp.generate_rmat(12, 16, seed=3)
    cases.append(("rmat12_k16", off, nbr, None, None, 16))
    off, nbr = gp.generate_rgg2d(1 << 13, (8.0 / (3.14159265 * (1 << 13))) ** 0.5, seed=5)
    cases.append(("rgg13_k8", off, nbr, None, None, 8))
    off, nbr, ew, vw = gp.generate_knn3d(3000, 6, seed=2)
    cases.append(("knn3000_k16_weighted", off, nbr, ew, vw, 16))
    off, nbr = gp.generate_rmat(15, 16, seed=1)
    cases.append(("rmat15_k64", off, nbr, None, None, 64))
    off, nbr, ew, vw = gp.generate_knn3d(20000, 10, seed=4)
    cases.append(("knn20000_k128_weighted", off, nbr, ew, vw, 128))
    off, nbr = gp.generate_rmat(11, 8, seed=9)
    cases.append(("rmat11_k2", off, nbr, None, None, 2))
    # Rows longer than 4096 entries (17 of them, up to 9647) run in the
    # heavy tier on the device.
    off, nbr = gp.generate_rmat(16, 16, seed=1)
    cases.append(("rmat16_k64_heavy_rows", off, nbr, None, None, 64))
    off, nbr, ew, vw = graphs.hub_ring(12000, 6000, seed=4)
    cases.append(("hub12000_k8_skewed_weights", off, nbr, ew, vw, 8))
    # A hub row of 19,999 entries: above the 12,288-entry limit of the
    # one-block LDS hash, so its affinity goes through the global records.
    off, nbr, ew, vw = graphs.hub_ring(20000, 10000, seed=6)
    cases.append(("hub20000_k8_row_beyond_block_hash", off, nbr, ew, vw, 8))
    # Row weights above 2^32 at level 0 and after contraction: 64-bit part
    # tables, the sort-based affinity and the plain record path.
    off, nbr, ew, vw = graphs.heavy_weights(3000, 9000, seed=12)
    cases.append(("heavy3000_k4_row_weight_beyond_u32", off, nbr, ew, vw, 4))
    return cases


@pytest.fixture(scope="module")
def cases(gp):
    return _cases(gp)


def _compare_hierarchy(gp, og, ctx, name, off, nbr, ew, vw, k):
    _say("case %s: coarsen n=%d m=%d k=%d" % (name, len(off) - 1, len(nbr), k))
    g = gp.Graph(ctx, off, nbr, ew, vw)
    h = gp.Hierarchy(ctx, g, gp.coarsen_options(k))
    oh = og.Hierarchy(off, nbr, ew, vw, og.coarsen_opts(k))
    assert h.levels == oh.levels, (name, h.levels, oh.levels)
    for lvl in range(h.levels):
        a = h.graph_at(lvl)
        b = oh.graph(lvl)
        for x, y, what in zip(a, b, ("offsets", "neighbors", "edge_weights", "vertex_weights")):
            assert np.array_equal(x, y), (name, lvl, what)
        if lvl + 1 < h.levels:
            assert np.array_equal(h.map(lvl), oh.map(lvl)), (name, lvl, "map")
        gi, oi = h.info(lvl), oh.info(lvl)
        assert gi["max_degree"] == oi["max_degree"] and gi["has_heavy"] == oi["has_heavy"], (name, lvl, gi, oi)
    return g, h, oh


def test_coarsen_and_uncoarsen_bit_exact(gp, og, ctx, cases):
    for name, off, nbr, ew, vw, k in cases:
        g, h, oh = _compare_hierarchy(gp, og, ctx, name, off, nbr, ew, vw, k)
        coff, cnbr, cew, cvw = oh.graph(oh.levels - 1)
        init = gp.initial_partition_host(coff, cnbr, cew, cvw, k)
        _say("case %s: uncoarsen from %d coarse vertices" % (name, len(coff) - 1))
        lab, cut, loads, st = h.uncoarsen(init, gp.refine_options(k))
        olab, ocut, oloads, ost = oh.uncoarsen(init, og.refine_opts(k))
        assert cut == ocut, (name, cut, ocut)
        assert np.array_equal(loads, oloads), name
        assert np.array_equal(lab, olab), (name, int((lab != olab).sum()))
        for a, b in zip(st, ost):
            for key in ("n", "refined", "initial_repair_failed", "rounds", "initial_cut", "final_cut",
                        "lp_moves", "repair_moves", "group_moves", "repair_iterations"):
                assert a[key] == b[key], (name, key, a, b)
        # independent recount and strict capacity
        assert og.cut(off, nbr, ew, lab) == cut
        w = np.ones(len(off) - 1, np.uint64) if vw is None else vw.astype(np.uint64)
        assert np.array_equal(np.bincount(lab, weights=w, minlength=k).astype(np.uint64), loads)
        assert int(loads.max()) <= og.capacity(int(w.sum()), k)
        h.close()
        g.close()
        _say("case %s: ok cut=%d" % (name, cut))


def test_full_partition_matches_stages(gp, og, ctx):
    off, nbr = gp.generate_rmat(13, 16, seed=7)
    k = 32
    _say("case rmat13_k32: gp_partition")
    g = gp.Graph(ctx, off, nbr)
    labels, st = g.partition(k)
    oh = og.Hierarchy(off, nbr, None, None, og.coarsen_opts(k))
    coff, cnbr, cew, cvw = oh.graph(oh.levels - 1)
    init = gp.initial_partition_host(coff, cnbr, cew, cvw, k)
    olab, ocut, _, _ = oh.uncoarsen(init, og.refine_opts(k))
    assert np.array_equal(labels, olab)
    assert st["final_cut"] == ocut
    assert st["max_part_weight"] <= st["capacity"]
    g.close()


def test_refine_partition_single_level(gp, og, ctx):
    off, nbr, ew = graphs.random_graph(3000, 12000, seed=11, max_weight=9)
    k = 8
    rng = np.random.default_rng(4)
    init = rng.integers(0, k, len(off) - 1).astype(np.int32)
    _say("case random3000_refine")
    g = gp.Graph(ctx, off, nbr, ew, None)
    lab, st = g.refine(init, gp.refine_options(k))
    olab, ost = og.refine_graph(off, nbr, ew, None, init, og.refine_opts(k))
    assert np.array_equal(lab, olab)
    assert st["final_cut"] == ost["final_cut"]
    g.close()


def test_coarse_edge_weight_overflow_is_an_error_on_both(gp, og, ctx):
    # At k=16 two heavy edges merge into one coarse edge above 2^32 - 1: the
    # oracle and the device both refuse (coarsen.cu throws likewise).
    off, nbr, ew, vw = graphs.heavy_weights(3000, 9000, seed=12)
    _say("case heavy3000_k16: overflow")
    with pytest.raises(RuntimeError, match="coarse edge weight exceeds uint32"):
        og.Hierarchy(off, nbr, ew, vw, og.coarsen_opts(16))
    g = gp.Graph(ctx, off, nbr, ew, vw)
    try:
        with pytest.raises(RuntimeError, match="coarse edge weight exceeds uint32"):
            gp.Hierarchy(ctx, g, gp.coarsen_options(16))
    finally:
        g.close()


def test_edge_cut_of_device_labels(gp, og, ctx):
    # gp_edge_cut on the result's device labels counts the same cut and
    # loads as the refinement and as a host recount; labels outside the
    # part range are refused.
    cases = [("grid64_k4", graphs.grid(64, 64) + (None, None), 4)]
    off, nbr, ew, vw = gp.generate_knn3d(5000, 8, seed=3)
    cases.append(("knn5000_k16_weighted", (off, nbr, ew, vw), 16))
    off, nbr = gp.generate_rmat(14, 16, seed=2)
    cases.append(("rmat14_k64", (off, nbr, None, None), 64))
    for name, (off, nbr, ew, vw), k in cases:
        _say("case %s: edge cut of device labels" % name)
        g = gp.Graph(ctx, off, nbr, ew, vw)
        h = gp.Hierarchy(ctx, g, gp.coarsen_options(k))
        try:
            coff, cnbr, cew, cvw = h.graph_at(h.levels - 1)
            init = gp.initial_partition_host(coff, cnbr, cew, cvw, k)
            lab, cut, loads, _, (ecut, eloads) = h.uncoarsen(init, gp.refine_options(k), evaluate_parts=k)
            assert ecut == cut == og.cut(off, nbr, ew, lab), name
            assert np.array_equal(eloads, loads), name
            with pytest.raises(RuntimeError, match="labels outside"):
                h.uncoarsen(init, gp.refine_options(k), evaluate_parts=2)
        finally:
            h.close()
            g.close()


def test_strict_verify_recounts_and_changes_nothing(gp, ctx, cases):
    # strict_verify (the CLI's --verify; include/refine.hpp:17) recounts the
    # cut and the loads after every projection and every committed batch of
    # LP, repair and exchange moves (refine.cu:641-659, incremental.cuh:846-853)
    # and must leave the result bit-identical.
    picked = {"grid256_k8", "knn3000_k16_weighted", "hub12000_k8_skewed_weights", "rmat16_k64_heavy_rows"}
    for name, off, nbr, ew, vw, k in cases:
        if name not in picked:
            continue
        _say("case %s: strict verification" % name)
        g = gp.Graph(ctx, off, nbr, ew, vw)
        h = gp.Hierarchy(ctx, g, gp.coarsen_options(k))
        try:
            coff, cnbr, cew, cvw = h.graph_at(h.levels - 1)
            init = gp.initial_partition_host(coff, cnbr, cew, cvw, k)
            lab, cut, loads, st = h.uncoarsen(init, gp.refine_options(k))
            slab, scut, sloads, sst = h.uncoarsen(init, gp.refine_options(k, strict_verify=1))
        finally:
            h.close()
            g.close()
        assert np.array_equal(lab, slab) and cut == scut and np.array_equal(loads, sloads), name
        assert all(s["strict_checks"] == 0 for s in st), name
        for lvl, s in enumerate(sst):
            moved = s["lp_moves"] + s["repair_moves"] + s["group_moves"]
            if lvl > 0:
                assert s["strict_checks"] >= 1, (name, lvl, s)  # the projection check
            if s["refined"] and not s["initial_repair_failed"]:
                assert s["strict_checks"] >= 1 + (1 if moved else 0), (name, lvl, s)
        _say("case %s: ok, %d recounts" % (name, sum(s["strict_checks"] for s in sst)))
