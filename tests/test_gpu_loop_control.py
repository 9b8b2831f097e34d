"""LP loop control: the round loop decided on the device (the default), the
same with each round issued before the previous one is seen to end
(GP_LOOP_AHEAD=1), and the loop decided on the host after a counter
read-back per round (the reference's option gpu_lp_loop = false, or
GP_DEVICE_LOOP=0 for a whole context) give the same partitions -- labels,
cut, loads and every level's statistics -- on graphs whose levels run the
small-level (long double) and graph-mode stale rules, repairs in the middle
of the loop, and priced proposals.  Each mode is a context created with its
environment setting (read when the context is created)."""
import os

import numpy as np
import pytest

from tests import graphs

pytestmark = pytest.mark.gpu

# (name, environment of the context, refinement option overrides)
MODES = (("device", {}, {}), ("ahead", {"GP_LOOP_AHEAD": "1"}, {}), ("host", {"GP_DEVICE_LOOP": "0"}, {}),
         ("option", {}, {"gpu_lp_loop": 0}))
KEYS = ("n", "refined", "initial_repair_failed", "rounds", "initial_cut", "final_cut", "lp_moves", "repair_moves",
        "group_moves", "repair_iterations")


def _context(gp, env):
    saved = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return gp.Context(0)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _cases(gp):
    off, nbr = graphs.grid(32, 32)
    yield "grid32_k4", off, nbr, None, None, 4
    off, nbr = gp.generate_grid(256, 256)
    yield "grid256_k8", off, nbr, None, None, 8
    off, nbr = gp.generate_rmat(13, 16, seed=7)
    yield "rmat13_k32", off, nbr, None, None, 32
    off, nbr, ew, vw = gp.generate_knn3d(3000, 6, seed=2)
    yield "knn3000_k16_weighted", off, nbr, ew, vw, 16
    off, nbr, ew, vw = graphs.hub_ring(12000, 6000, seed=4)
    yield "hub12000_k8_skewed_weights", off, nbr, ew, vw, 8


def test_device_loop_equals_host_loop(gp):
    ctxs = [(name, _context(gp, env), ropts) for name, env, ropts in MODES]
    try:
        for case, off, nbr, ew, vw, k in _cases(gp):
            print("case %s: n=%d k=%d" % (case, len(off) - 1, k), flush=True)
            results = []
            for mode, c, ropts in ctxs:
                g = gp.Graph(c, off, nbr, ew, vw)
                h = gp.Hierarchy(c, g, gp.coarsen_options(k))
                coff, cnbr, cew, cvw = h.graph_at(h.levels - 1)
                init = gp.initial_partition_host(coff, cnbr, cew, cvw, k)
                lab, cut, loads, st = h.uncoarsen(init, gp.refine_options(k, **ropts))
                results.append((mode, lab, cut, loads, st))
                h.close()
                g.close()
            _, lab0, cut0, loads0, st0 = results[0]
            assert sum(s["rounds"] for s in st0) > 0, case
            for mode, lab, cut, loads, st in results[1:]:
                assert cut == cut0, (case, mode, cut, cut0)
                assert np.array_equal(loads, loads0), (case, mode)
                assert np.array_equal(lab, lab0), (case, mode, int((lab != lab0).sum()))
                assert len(st) == len(st0), (case, mode)
                for a, b in zip(st, st0):
                    for key in KEYS:
                        assert a[key] == b[key], (case, mode, key, a, b)
    finally:
        for _, c, _ in ctxs:
            c.close()
