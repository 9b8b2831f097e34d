"""Batched and served forms: every result equals the same graph partitioned
alone (bit-exact; deterministic mode is the only mode)."""
import numpy as np
import pytest

from tests import graphs

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _mix(gp):
    items = []
    rng = np.random.default_rng(5)
    for i in range(10):
        kind = i % 3
        k = int(rng.choice([4, 8, 16, 32, 64]))
        if kind == 0:
            side = int(rng.integers(32, 160))
            off, nbr = gp.generate_grid(side, side)
            items.append(("grid%d_k%d" % (side, k), off, nbr, None, None, k))
        elif kind == 1:
            scale = int(rng.integers(10, 15))
            off, nbr = gp.generate_rmat(scale, int(rng.integers(8, 17)), seed=int(rng.integers(1, 1000)))
            items.append(("rmat%d_k%d" % (scale, k), off, nbr, None, None, k))
        else:
            n = 1 << int(rng.integers(10, 15))
            off, nbr = gp.generate_rgg2d(n, (8.0 / (np.pi * n)) ** 0.5, seed=int(rng.integers(1, 1000)))
            items.append(("rgg%d_k%d" % (n, k), off, nbr, None, None, k))
    off, nbr, ew, vw = gp.generate_knn3d(4000, 6, seed=2)
    items.append(("knn4000_k16_weighted", off, nbr, ew, vw, 16))
    return items


def test_batch_equals_alone(gp, ctx):
    mix = _mix(gp)
    alone = []
    for name, off, nbr, ew, vw, k in mix:
        g = gp.Graph(ctx, off, nbr, ew, vw)
        lab, st = g.partition(k)
        g.close()
        alone.append((lab, st["final_cut"]))
    print("batch of %d graphs" % len(mix), flush=True)
    items = [gp.BatchItem(off, nbr, k, ew=ew, vw=vw) for _, off, nbr, ew, vw, k in mix]
    stats = gp.partition_batch(items, workers=4)
    for (name, off, nbr, ew, vw, k), it, st, (lab, cut) in zip(mix, items, stats, alone):
        assert np.array_equal(it.labels, lab), name
        assert st["final_cut"] == cut, name


def test_server_mixed_order_and_warm_start(gp, ctx, og):
    mix = _mix(gp)[:6]
    srv = gp.Server(workers=3, hbm_budget=1 << 30)
    try:
        items = [gp.BatchItem(off, nbr, k, ew=ew, vw=vw) for _, off, nbr, ew, vw, k in mix]
        tickets = [srv.submit(it) for it in reversed(items)]
        for t in tickets:
            srv.wait(t)
        for (name, off, nbr, ew, vw, k), it in zip(mix, items):
            g = gp.Graph(ctx, off, nbr, ew, vw)
            lab, _ = g.partition(k)
            g.close()
            assert np.array_equal(it.labels, lab), name
        # Warm start: refine prior labels on a perturbed graph, as
        # refine_partition would alone.
        name, off, nbr, ew, vw, k = mix[0]
        rng = np.random.default_rng(3)
        n = len(off) - 1
        extra = set()
        while len(extra) < max(1, len(nbr) // 200):
            u, v = (int(x) for x in rng.integers(0, n, 2))
            if u != v:
                extra.add((min(u, v), max(u, v)))
        src = np.repeat(np.arange(n), np.diff(off))
        edges = set(zip(src.tolist(), nbr.tolist()))
        pairs = sorted({(a, b) for a, b in edges if a < b} | extra)
        off2, nbr2, _ = graphs.csr_from_edges(n, pairs)
        warm = items[0].labels.copy()
        it = gp.BatchItem(off2, nbr2, k, warm=warm)
        srv.wait(srv.submit(it))
        g = gp.Graph(ctx, off2, nbr2)
        ref, st = g.refine(warm, gp.refine_options(k))
        g.close()
        assert np.array_equal(it.labels, ref)
        assert og.cut(off2, nbr2, None, it.labels) == st["final_cut"]
    finally:
        srv.close()


def test_apply_delta_matches_host(gp, ctx):
    # gp_graph_apply_delta against its host restatement (tests/workloads.py),
    # on a unit grid (also with weights that force an explicit weight
    # array), a weighted kNN graph and an R-MAT graph with hub rows; edge
    # cases: an absent deletion, a present insertion, a repeated insertion,
    # a deletion and re-insertion of one edge.
    from tests import graphs, workloads
    rng = np.random.default_rng(9)
    off, nbr = graphs.grid(40, 40)
    cases = [("grid40", off, nbr, None, None, False), ("grid40_weighted_inserts", off, nbr, None, None, True)]
    o, nb, ew, vw = gp.generate_knn3d(3000, 8, seed=4)
    cases.append(("knn3000_weighted", o, nb, ew, vw, True))
    o, nb = gp.generate_rmat(12, 16, seed=2)
    cases.append(("rmat12", o, nb, None, None, False))
    for name, off, nbr, ew, vw, with_w in cases:
        _say("case %s: edge delta" % name)
        u, v, ins = workloads.random_delta(off, nbr, 0.02, rng)
        a = int(nbr[off[5]])
        row5 = set(nbr[off[5]:off[6]].tolist())
        b = next(x for x in range(len(off) - 1) if x != 5 and x not in row5)  # not adjacent to 5
        u = np.concatenate([u, [5, 5, 11, 11, a, a]]).astype(np.int32)
        v = np.concatenate([v, [b, a, 17, 17, 5, 5]]).astype(np.int32)
        ins = np.concatenate([ins, [0, 1, 1, 1, 0, 1]]).astype(np.uint8)
        w = rng.integers(1, 30, len(u)).astype(np.uint32) if with_w else None
        want_off, want_nbr, want_ew = workloads.apply_delta(off, nbr, ew, u, v, ins, w)
        g = gp.Graph(ctx, off, nbr, ew, vw)
        try:
            m = gp.apply_delta(ctx, g.h, u, v, ins, w)
            got_off, got_nbr, got_ew, _ = gp.download_graph(ctx, g.h)
            assert m == len(want_nbr)
            assert np.array_equal(got_off, want_off) and np.array_equal(got_nbr, want_nbr), name
            assert np.array_equal(got_ew, want_ew if want_ew is not None else np.ones(m, np.uint32)), name
            with pytest.raises(RuntimeError, match="invalid vertex or a self loop"):
                gp.apply_delta(ctx, g.h, np.array([3], np.int32), np.array([3], np.int32), np.array([1], np.uint8))
            assert np.array_equal(gp.download_graph(ctx, g.h)[1], want_nbr)
        finally:
            g.close()


def test_request_shares_equal_alone(gp, ctx):
    """With request shares (GP_SHARE_ENTRIES, read when the server starts) a
    request of m CSR entries takes 1 + m / E of the workers' slots: large
    requests run beside fewer others.  Only the schedule changes, so every
    result equals the same graph partitioned alone."""
    import os
    mix = _mix(gp)
    alone = []
    for name, off, nbr, ew, vw, k in mix:
        g = gp.Graph(ctx, off, nbr, ew, vw)
        lab, st = g.partition(k)
        g.close()
        alone.append(lab)
    saved = os.environ.get("GP_SHARE_ENTRIES")
    os.environ["GP_SHARE_ENTRIES"] = "20000"
    try:
        items = [gp.BatchItem(off, nbr, k, ew=ew, vw=vw) for _, off, nbr, ew, vw, k in mix]
        gp.partition_batch(items, workers=4)
    finally:
        if saved is None:
            os.environ.pop("GP_SHARE_ENTRIES", None)
        else:
            os.environ["GP_SHARE_ENTRIES"] = saved
    assert max(len(nbr) for _, _, nbr, _, _, _ in mix) > 3 * 20000  # some requests take several slots
    for (name, *_), it, lab in zip(mix, items, alone):
        assert np.array_equal(it.labels, lab), name
