"""Full-size configurations, checked through size-independent properties.

The oracle is too slow at these sizes, so each run is checked by what holds
for every correct run: labels in range, every part load within the integer
capacity, the reported cut equal to an independent recount (numpy and the
gpart_verify_csr tool), bit-identical labels on a second run (deterministic
mode), and per-level contraction invariants (vertex and edge weight
conservation, canonical sorted symmetric rows without self loops).

Workloads: R-MAT scale 22, edge factor 16 (a=0.57, b=0.19, c=0.19), k=64,
eps=1.03; and the weighted 3-D 10-nearest-neighbour graph of 8M points
(edge weights 1..100, vertex weights 1..4), k=128, eps=1.03.
"""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
VERIFY = os.path.join(ROOT, "acc-gp_amd", "bin", "gpart_verify_csr")


def _say(msg):
    print(msg, flush=True)


def _cut(off, nbr, ew, lab):
    src = np.repeat(np.arange(len(off) - 1, dtype=np.int32), np.diff(off))
    cross = lab[src] != lab[nbr]
    w = np.ones(len(nbr), np.uint64) if ew is None else ew.astype(np.uint64)
    return int(w[cross].sum()) // 2


def _check_partition(gp, og, off, nbr, ew, vw, k, lab, st):
    n = len(off) - 1
    assert lab.shape == (n,) and lab.min() >= 0 and lab.max() < k
    w = np.ones(n, np.uint64) if vw is None else vw.astype(np.uint64)
    loads = np.bincount(lab, weights=w.astype(np.float64), minlength=k).astype(np.uint64)
    cap = og.capacity(int(w.sum()), k)
    assert st["capacity"] == cap
    assert int(loads.max()) == st["max_part_weight"] <= cap
    assert _cut(off, nbr, ew, lab) == st["final_cut"]
    return loads


def _check_contraction(fine, coarse, cmap):
    """Level invariants of one contraction step (fine -> coarse)."""
    off, nbr, ew, vw = fine
    coff, cnbr, cew, cvw = coarse
    nc = len(coff) - 1
    assert cmap.min() >= 0 and cmap.max() == nc - 1
    assert np.array_equal(np.bincount(cmap, weights=vw.astype(np.float64), minlength=nc).astype(np.uint64),
                          cvw.astype(np.uint64))
    src = np.repeat(np.arange(len(off) - 1, dtype=np.int32), np.diff(off))
    external = cmap[src] != cmap[nbr]
    assert int(ew.astype(np.uint64)[external].sum()) == int(cew.astype(np.uint64).sum())
    csrc = np.repeat(np.arange(nc, dtype=np.int64), np.diff(coff))
    assert not np.any(csrc == cnbr), "self loop in a coarse row"
    row_change = np.diff(csrc) != 0
    assert np.all(row_change | (np.diff(cnbr) > 0)), "coarse rows not strictly ascending"
    fwd = csrc * nc + cnbr
    rev = cnbr.astype(np.int64) * nc + csrc
    order = np.argsort(rev, kind="stable")
    assert np.array_equal(rev[order], fwd), "coarse graph not symmetric"
    assert np.array_equal(cew[order], cew), "coarse edge weights not symmetric"


def _verify_tool(gp, tmp_path, off, nbr, ew, vw, k, lab):
    ip, ix = str(tmp_path / "g.indptr"), str(tmp_path / "g.indices")
    gp.save_csr(ip, ix, off, nbr)
    part = str(tmp_path / "part.bin")
    lab.astype("<i4").tofile(part)
    cmd = [VERIFY, ip, ix, str(k), part, "binary"]
    if ew is not None:
        ewf, vwf = str(tmp_path / "g.ew"), str(tmp_path / "g.vw")
        ew.astype("<u4").tofile(ewf)
        vw.astype("<u4").tofile(vwf)
        cmd += ["--edge-weights", ewf, "--vertex-weights", vwf]
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    return json.loads(res.stdout)


def _full_size(gp, og, ctx, tmp_path, name, off, nbr, ew, vw, k, contraction_levels):
    _say("case %s: partition n=%d m=%d k=%d" % (name, len(off) - 1, len(nbr), k))
    g = gp.Graph(ctx, off, nbr, ew, vw)
    try:
        lab, st = g.partition(k)
        loads = _check_partition(gp, og, off, nbr, ew, vw, k, lab, st)
        _say("case %s: cut=%d levels=%d; second run" % (name, st["final_cut"], st["levels"]))
        lab2, st2 = g.partition(k)
        assert np.array_equal(lab, lab2) and st2["final_cut"] == st["final_cut"]
        _say("case %s: verify tool" % name)
        got = _verify_tool(gp, tmp_path, off, nbr, ew, vw, k, lab)
        assert got["cut"] == st["final_cut"]
        assert got["loads"] == loads.tolist()
        _say("case %s: contraction invariants" % name)
        h = gp.Hierarchy(ctx, g, gp.coarsen_options(k))
        try:
            assert h.levels == st["levels"]
            fine = h.graph_at(0)
            for lvl in range(min(contraction_levels, h.levels - 1)):
                coarse = h.graph_at(lvl + 1)
                _check_contraction(fine, coarse, h.map(lvl))
                assert len(coarse[0]) < len(fine[0])
                fine = coarse
        finally:
            h.close()
    finally:
        g.close()
    _say("case %s: ok" % name)


def test_rmat22_k64_full_size(gp, og, ctx, tmp_path):
    off, nbr = gp.generate_rmat(22, 16, seed=1)
    _full_size(gp, og, ctx, tmp_path, "rmat22_k64", off, nbr, None, None, 64, contraction_levels=3)


def test_knn8m_k128_weighted_full_size(gp, og, ctx, tmp_path):
    off, nbr, ew, vw = gp.generate_knn3d(1 << 23, 10, seed=2)
    _full_size(gp, og, ctx, tmp_path, "knn8m_k128_weighted", off, nbr, ew, vw, 128, contraction_levels=2)
