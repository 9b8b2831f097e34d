"""gp_edge_cut_noalloc (the allocation-free form of gp_edge_cut) against
gp_edge_cut and a host recount, through the C ABI: two calls in a row, the
second on a smaller input, bit-exact cut and loads.  The labels are the
device labels of partition results; the evaluations run on a second, fresh
context whose arena holds nothing, so any allocation from it would show in
its reserved bytes.
"""
import numpy as np
import pytest

from tests import graphs

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _host_counts(og, off, nbr, ew, vw, labels, k):
    cut = og.cut(off, nbr, ew, labels)
    w = np.ones(len(labels), np.uint64) if vw is None else vw.astype(np.uint64)
    loads = np.bincount(labels, weights=w, minlength=k).astype(np.uint64)
    return cut, loads


def _partition_and_evaluate(gp, ctx, off, nbr, ew, vw, k, evaluate):
    g = gp.Graph(ctx, off, nbr, ew, vw)
    h = gp.Hierarchy(ctx, g, gp.coarsen_options(k))
    try:
        coff, cnbr, cew, cvw = h.graph_at(h.levels - 1)
        init = gp.initial_partition_host(coff, cnbr, cew, cvw, k)
        lab, cut, loads, _, got = h.uncoarsen(init, gp.refine_options(k), evaluate=evaluate)
        return lab, cut, loads, got
    finally:
        h.close()
        g.close()


def test_noalloc_cut_equals_arena_cut_twice(gp, og, ctx):
    with gp.Context(0) as fresh:
        assert gp.reserved_bytes(fresh) == 0

        def both(k):
            def run(g_handle, device_labels):
                before = gp.reserved_bytes(fresh)
                a = gp.edge_cut_noalloc(fresh, g_handle, device_labels, k)
                after_noalloc = gp.reserved_bytes(fresh)
                b = gp.edge_cut(fresh, g_handle, device_labels, k)
                return a, b, before, after_noalloc
            return run

        _say("case rmat15 k=128: no-alloc evaluation on a fresh context")
        off1, nbr1 = gp.generate_rmat(15, 16, seed=4)
        lab, cut, loads, (a, b, before, after) = _partition_and_evaluate(gp, ctx, off1, nbr1, None, None, 128,
                                                                         both(128))
        assert before == 0 and after == 0, "the no-alloc evaluation took memory from the arena"
        assert gp.reserved_bytes(fresh) > 0  # the arena form allocates its partials
        want_cut, want_loads = _host_counts(og, off1, nbr1, None, None, lab, 128)
        assert a[0] == b[0] == cut == want_cut
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[1], want_loads) and np.array_equal(a[1], loads)

        _say("case knn3000 weighted k=7: second call, smaller input")
        off2, nbr2, ew2, vw2 = gp.generate_knn3d(3000, 6, seed=9)
        assert len(nbr2) < len(nbr1) and len(off2) < len(off1)
        lab, cut, loads, (a, b, before, after) = _partition_and_evaluate(gp, ctx, off2, nbr2, ew2, vw2, 7, both(7))
        assert after == before, "the no-alloc evaluation took memory from the arena"
        want_cut, want_loads = _host_counts(og, off2, nbr2, ew2, vw2, lab, 7)
        assert a[0] == b[0] == cut == want_cut
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[1], want_loads) and np.array_equal(a[1], loads)

        _say("case errors: parts above 128, labels outside [0, parts)")

        def refused(g_handle, device_labels):
            before = gp.reserved_bytes(fresh)
            errors = []
            for parts in (129, 2):
                try:
                    gp.edge_cut_noalloc(fresh, g_handle, device_labels, parts)
                    errors.append(None)
                except RuntimeError as e:
                    errors.append(str(e))
            return errors, before, gp.reserved_bytes(fresh)

        _, _, _, (errors, before, after) = _partition_and_evaluate(gp, ctx, off2, nbr2, ew2, vw2, 7, refused)
        assert errors[0] is not None and "parts must be in" in errors[0]
        assert errors[1] is not None and "labels outside" in errors[1]
        assert after == before

        _say("case grid40 k=4: after the refused calls")
        off3, nbr3 = graphs.grid(40, 40)
        lab, cut, loads, (a, b, _, _) = _partition_and_evaluate(gp, ctx, off3, nbr3, None, None, 4, both(4))
        want_cut, want_loads = _host_counts(og, off3, nbr3, None, None, lab, 4)
        assert a[0] == b[0] == want_cut and np.array_equal(a[1], want_loads)
