"""Full-size parity: the device run of R-MAT scale 22 at k=64 and of the
weighted 8M-point 10-nearest-neighbour graph at k=128 against the oracle's
run of the same schedule, stored as SHA-256 digests in tests/golden/ (the
oracle takes minutes at these sizes; tests/golden/make_golden.py wrote the
fixtures).  Every level's offsets, neighbours, weights, map and metadata,
the host initial partition, the refined labels, cut, loads and per-level
statistics must match bit for bit, through the C ABI.
"""
import json
import os

import numpy as np
import pytest

from tests.golden.make_golden import WORKLOADS, digest, generate

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STAT_KEYS = ("n", "m", "refined", "initial_repair_failed", "rounds", "initial_cut", "final_cut", "lp_moves",
             "repair_moves", "group_moves", "repair_iterations")


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("name", sorted(WORKLOADS))
def test_full_size_bit_exact(gp, ctx, name):
    path = os.path.join(HERE, name + ".json")
    fx = json.load(open(path))
    _say("case %s: generate" % name)
    off, nbr, ew, vw = generate(gp, fx["spec"])
    assert digest(off, nbr, ew, vw) == fx["input"], "generator output changed"
    k = fx["spec"]["parts"]
    g = gp.Graph(ctx, off, nbr, ew, vw)
    h = None
    try:
        _say("case %s: coarsen" % name)
        h = gp.Hierarchy(ctx, g, gp.coarsen_options(k))
        assert h.levels == len(fx["levels"])
        for lvl, want in enumerate(fx["levels"]):
            goff, gnbr, gew, gvw = h.graph_at(lvl)
            info = h.info(lvl)
            assert len(goff) - 1 == want["n"] and len(gnbr) == want["m"], (name, lvl)
            for key in ("max_degree", "has_heavy", "incident_bound"):
                assert info[key] == want[key], (name, lvl, key)
            for key, arr in (("offsets", goff), ("neighbors", gnbr), ("edge_weights", gew),
                             ("vertex_weights", gvw)):
                assert digest(arr) == want[key], (name, lvl, key)
            if "map" in want:
                assert digest(h.map(lvl)) == want["map"], (name, lvl, "map")
        coff, cnbr, cew, cvw = h.graph_at(h.levels - 1)
        init = gp.initial_partition_host(coff, cnbr, cew, cvw, k)
        assert digest(init) == fx["initial_partition"]
        _say("case %s: uncoarsen" % name)
        lab, cut, loads, stats = h.uncoarsen(init, gp.refine_options(k, fx["spec"]["imbalance"]))
        assert cut == fx["cut"], (name, cut, fx["cut"])
        assert [int(x) for x in loads] == fx["loads"]
        assert digest(lab) == fx["labels"]
        for got, want in zip(stats, fx["level_stats"]):
            for key in STAT_KEYS:
                assert got[key] == want[key], (name, key, got, want)
        h.close()
        h = None
        _say("case %s: whole pipeline" % name)
        lab2, st = g.partition(k, fx["spec"]["imbalance"], 0)
        assert digest(lab2) == fx["labels"] and st["final_cut"] == fx["cut"]
    finally:
        if h is not None:
            h.close()
        g.close()
    _say("case %s: ok cut=%d" % (name, fx["cut"]))
