"""The CPU oracle pinned against known answers, brute force and invariants.

The reference ships no tests or golden vectors for this path; the pinned
answers are the worked examples of docs/GPU_ARCHITECTURE_ANALYSIS.md
(admission: lines 847-892; capacity-complementary exchange: 1228-1247), the
reference's own partition_capacity() and gpart_verify_csr built from its
sources into oracle/_ref (when present), brute-force contraction, and the
invariants the reference checks in strict mode (coarsen.cu:1692-1758,
refine.cu:641-659, incremental.cuh:846-853,1045-1048).
"""
import fractions
import math
import os
import subprocess

import numpy as np
import pytest

from tests import graphs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_CAPACITY = os.path.join(ROOT, "oracle", "_ref", "capacity_ref")


# ---------------------------------------------------------------- capacity
def _capacity_closed_form(total, parts, imbalance):
    # floor(W * eps / k) with the double eps taken exactly; a capacity below
    # ceil(W / k) cannot hold the total weight and is an error (0 here).
    cap = math.floor(fractions.Fraction(total) * fractions.Fraction(imbalance) / parts)
    return cap if cap >= -(-total // parts) else 0


CAPACITY_CASES = [(65536, 8, 1.03), (4194304, 64, 1.03), (20, 2, 1.0), (7, 3, 1.03), (1000003, 128, 1.05),
                  (8 * 2 ** 20 * 3, 128, 1.03), (12345678901, 64, 1.03)]


def test_capacity_closed_form(og):
    for total, parts, eps in CAPACITY_CASES:
        assert og.capacity(total, parts, eps) == _capacity_closed_form(total, parts, eps), (total, parts, eps)
    # Published anchors: grid 256x256 k=8 and R-MAT scale 22 k=64.
    assert og.capacity(65536, 8, 1.03) == 8437
    assert og.capacity(4194304, 64, 1.03) == 67502


@pytest.mark.skipif(not os.path.exists(REF_CAPACITY), reason="reference capacity checker not built")
def test_capacity_matches_reference_build(og):
    text = "".join("%d %d %r\n" % c for c in CAPACITY_CASES)
    out = subprocess.run([REF_CAPACITY], input=text, capture_output=True, text=True, check=True).stdout.split()
    assert [0 if x == "error" else int(x) for x in out] == [og.capacity(*c) for c in CAPACITY_CASES]


# --------------------------------------------------------------- admission
def test_admission_mixed_example(og):
    # C=4; receivers A (weight 2, id 0) and B (weight 1, id 1).  Requests
    # u (w=1 -> A, regret 1), v (w=2 -> A, regret 6), z (w=1 -> B, regret 4):
    # v joins A, z joins B, u is rejected.
    acc = og.admission_batch(targets=[0, 0, 1], keys=[1, 6, 4], ties=[0, 0, 0], vertices=[10, 11, 12],
                             weights=[1, 2, 1], bases=[2, 2, 1], cap=4)
    assert acc.tolist() == [False, True, True]


def test_admission_regret_order_example(og):
    # Two weight-1 vertices compete for the one free slot of A (cap 3,
    # weight 2); v has the larger regret and wins.
    acc = og.admission_batch(targets=[0, 0], keys=[1, 6], ties=[0, 0], vertices=[1, 2], weights=[1, 1],
                             bases=[2, 2], cap=3)
    assert acc.tolist() == [False, True]


def test_admission_is_a_prefix_not_bin_packing(og):
    # Room 4, request weights 2, 3, 1 in order: prefixes 2, 5, 6 admit only
    # the first; the third is not packed into the remaining room.
    acc = og.admission_batch(targets=[0, 0, 0], keys=[3, 2, 1], ties=[0, 0, 0], vertices=[5, 6, 7],
                             weights=[2, 3, 1], bases=[0, 0, 0], cap=4)
    assert acc.tolist() == [True, False, False]


def test_admission_tie_and_vertex_order(og):
    # Equal regret: larger tie first, then smaller vertex id.
    acc = og.admission_batch(targets=[3, 3, 3], keys=[5, 5, 5], ties=[1, 9, 9], vertices=[4, 8, 2],
                             weights=[1, 1, 1], bases=[0, 0, 0], cap=2)
    assert acc.tolist() == [False, True, True]
    acc = og.admission_batch(targets=[3, 3, 3], keys=[5, 5, 5], ties=[9, 9, 9], vertices=[4, 8, 2],
                             weights=[1, 1, 1], bases=[0, 0, 0], cap=2)
    assert acc.tolist() == [True, False, True]


# ---------------------------------------------------------------- exchange
def _exchange_example():
    # a:8, u1:1, u2:1 in A; b:8, v:2 in B; C = 10 (k=2, eps=1.0).
    a, u1, u2, b, v = range(5)
    edges = [(u1, a), (u1, b), (u2, a), (u2, b), (v, a), (v, b), (u1, u2), (u1, v)]
    weights = [1, 4, 1, 3, 4, 1, 2, 1]
    off, nbr, ew = graphs.csr_from_edges(5, edges, weights)
    vw = np.array([8, 1, 1, 8, 2], np.uint32)
    labels = np.array([0, 0, 0, 1, 1], np.int32)
    return off, nbr, ew, vw, labels


def test_exchange_worked_example(og):
    off, nbr, ew, vw, labels = _exchange_example()
    assert og.capacity(20, 2, 1.0) == 10
    assert og.cut(off, nbr, ew, labels) == 12
    out, st = og.refine_graph(off, nbr, ew, vw, labels, og.refine_opts(2, 1.0, exchange_only=1))
    # S = {u1, u2} moves to B and T = {v} to A: G(S, T) = 6 + 4 - 2 = 8.
    assert out.tolist() == [0, 1, 1, 1, 0]
    assert st["initial_cut"] == 12 and st["final_cut"] == 4 and st["group_moves"] == 3
    assert og.cut(off, nbr, ew, out) == 4


def test_joint_gain_closed_form_by_recount(og):
    # G(S,T) = sum g0 + 2 w(E[S]) + 2 w(E[T]) - 2 w(S,T) on the worked example.
    off, nbr, ew, vw, labels = _exchange_example()
    def cut_of(lab):
        return og.cut(off, nbr, ew, np.array(lab, np.int32))
    def g0(x):
        lab = labels.copy(); lab[x] = 1 - lab[x]
        return cut_of(labels) - cut_of(lab)
    S, T = [1, 2], [4]
    w = {}
    for x in range(5):
        for e in range(off[x], off[x + 1]):
            w[(x, int(nbr[e]))] = int(ew[e])
    inner = lambda A: sum(w.get((p, q), 0) for i, p in enumerate(A) for q in A[i + 1:])
    cross = sum(w.get((p, q), 0) for p in S for q in T)
    G = sum(g0(x) for x in S + T) + 2 * inner(S) + 2 * inner(T) - 2 * cross
    lab = labels.copy(); lab[S] = 1; lab[T] = 0
    assert G == 8 == cut_of(labels) - cut_of(lab)


# ------------------------------------------------------------- contraction
def _brute_contract(off, nbr, ew, cmap, nc):
    acc = {}
    for v in range(len(off) - 1):
        for e in range(off[v], off[v + 1]):
            a, b = int(cmap[v]), int(cmap[nbr[e]])
            if a != b:
                acc[(a, b)] = acc.get((a, b), 0) + (1 if ew is None else int(ew[e]))
    coff = [0]
    cnbr, cew = [], []
    for a in range(nc):
        row = sorted((b, s) for (x, b), s in acc.items() if x == a)
        cnbr += [b for b, _ in row]
        cew += [s for _, s in row]
        coff.append(len(cnbr))
    return np.array(coff, np.int32), np.array(cnbr, np.int32), np.array(cew, np.uint32)


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_contraction_brute_force(og, seed):
    rng = np.random.default_rng(seed)
    n = 300
    off, nbr, ew = graphs.random_graph(n, 1200, seed=seed, max_weight=7)
    nc = int(rng.integers(5, 60))
    cmap = rng.integers(0, nc, n).astype(np.int32)
    cmap[:nc] = np.arange(nc)  # every cluster non-empty
    got = og.contract(off, nbr, ew, cmap, nc)
    want = _brute_contract(off, nbr, ew, cmap, nc)
    for x, y in zip(got, want):
        assert np.array_equal(x, y)


# --------------------------------------------------------- hierarchy checks
def _check_csr(off, nbr, ew):
    n = len(off) - 1
    assert off[0] == 0 and off[-1] == len(nbr) and np.all(np.diff(off) >= 0)
    w = {}
    for v in range(n):
        row = nbr[off[v]:off[v + 1]]
        assert np.all(np.diff(row) > 0), "rows sorted and unique"
        assert not np.any(row == v), "no self loops"
        for e in range(off[v], off[v + 1]):
            w[(v, int(nbr[e]))] = int(ew[e])
    for (a, b), x in w.items():
        assert w.get((b, a)) == x, "symmetric weights"


def _hierarchy_cases(gp):
    off, nbr = graphs.grid(40, 40)
    yield "grid40_k8", off, nbr, None, None, 8
    off, nbr, ew = graphs.random_graph(2000, 9000, seed=5, max_weight=5)
    rng = np.random.default_rng(9)
    yield "random2000_k16_weighted", off, nbr, ew, rng.integers(1, 4, 2000).astype(np.uint32), 16
    off, nbr = gp.generate_rmat(11, 12, seed=4)
    yield "rmat11_k32", off, nbr, None, None, 32


def test_hierarchy_invariants(gp, og):
    for name, off, nbr, ew, vw, k in _hierarchy_cases(gp):
        h = og.Hierarchy(off, nbr, ew, vw, og.coarsen_opts(k))
        total = int(np.sum(vw, dtype=np.uint64)) if vw is not None else len(off) - 1
        assert h.levels >= 2, name
        rng = np.random.default_rng(1)
        for lvl in range(h.levels):
            coff, cnbr, cew, cvw = h.graph(lvl)
            _check_csr(coff, cnbr, cew)
            assert int(cvw.astype(np.uint64).sum()) == total, (name, lvl, "weight conservation")
            if lvl + 1 < h.levels:
                cmap = h.map(lvl)
                noff, nnbr, new_, nvw = h.graph(lvl + 1)
                nc = len(noff) - 1
                assert cmap.min() >= 0 and cmap.max() == nc - 1
                assert np.array_equal(np.bincount(cmap, weights=cvw, minlength=nc).astype(np.uint32), nvw)
                got = og.contract(coff, cnbr, cew, cmap, nc)
                for x, y in zip(got, (noff, nnbr, new_)):
                    assert np.array_equal(x, y), (name, lvl, "contraction")
                # projection preserves the cut (coarsen.cu:1692-1758)
                lab = rng.integers(0, 23, nc).astype(np.int32)
                assert og.cut(noff, nnbr, new_, lab) == og.cut(coff, cnbr, cew, lab[cmap]), (name, lvl)


def test_uncoarsen_postconditions(gp, og):
    for name, off, nbr, ew, vw, k in _hierarchy_cases(gp):
        h = og.Hierarchy(off, nbr, ew, vw, og.coarsen_opts(k))
        coff, cnbr, cew, cvw = h.graph(h.levels - 1)
        init = gp.initial_partition_host(coff, cnbr, cew, cvw, k)
        lab, cut, loads, stats = h.uncoarsen(init, og.refine_opts(k))
        w = np.ones(len(off) - 1, np.uint64) if vw is None else vw.astype(np.uint64)
        total = int(w.sum())
        assert og.cut(off, nbr, ew, lab) == cut, name
        assert np.array_equal(np.bincount(lab, weights=w, minlength=k).astype(np.uint64), loads), name
        assert int(loads.max()) <= og.capacity(total, k), name
        for st in stats:
            assert st["final_cut"] <= st["initial_cut"] or st["initial_repair_failed"], (name, st)


def test_refinement_cut_is_exact_on_random_starts(og):
    # The incremental cut must equal a full recount (incremental.cuh:846-853).
    for seed in range(4):
        off, nbr, ew = graphs.random_graph(400, 1600, seed=seed, max_weight=4)
        k = 4
        lab = np.random.default_rng(seed).integers(0, k, 400).astype(np.int32)
        out, st = og.refine_graph(off, nbr, ew, None, lab, og.refine_opts(k))
        assert og.cut(off, nbr, ew, out) == st["final_cut"]
        assert np.bincount(out, minlength=k).max() <= og.capacity(400, k)
        assert st["final_cut"] <= og.cut(off, nbr, ew, lab)


def test_schedule_stops_at_cutoff(gp, og):
    # k=64 gives cutoff max(32, beta*k) = 128 with beta = max(2, ceil(32/k)) = 2;
    # k=8 gives beta = max(64, ceil(512/8)) = 64 and cutoff 512
    # (app/partition_main.cu:75-78, src/coarsen.cu:1860-1938).
    off, nbr = gp.generate_grid(128, 128)
    for k, cutoff in ((64, 128), (8, 512)):
        h = og.Hierarchy(off, nbr, None, None, og.coarsen_opts(k))
        sizes = [h.info(l)["n"] for l in range(h.levels)]
        # every coarsened level started above the cutoff and shrank by the
        # stop ratio 0.90 at least; at most 24 coarsening steps
        assert all(s > cutoff for s in sizes[:-1]), (k, sizes)
        assert all(b <= 0.90 * a for a, b in zip(sizes, sizes[1:])), (k, sizes)
        assert h.levels <= 25


def test_oracle_threads_do_not_change_results(gp, og):
    # The OpenMP loops of the oracle write per-index outputs and integer
    # sums: one thread and all threads give the same hierarchy, labels,
    # loads and statistics (weighted and unit graphs, a hub row, k up to 64).
    cases = [(gp.generate_rmat(13, 16, seed=5) + (None, None), 64), (gp.generate_knn3d(6000, 8, seed=7), 16),
             (graphs.hub_ring(9000, 4000, seed=3), 8)]
    full = og.get_threads()
    try:
        for (off, nbr, ew, vw), k in cases:
            runs = []
            for threads in (1, max(full, 4)):
                og.set_threads(threads)
                h = og.Hierarchy(off, nbr, ew, vw, og.coarsen_opts(k))
                levels = [h.graph(lvl) for lvl in range(h.levels)]
                maps = [h.map(lvl) for lvl in range(h.levels - 1)]
                coff, cnbr, cew, cvw = levels[-1]
                init = (np.arange(len(coff) - 1) % k).astype(np.int32)
                lab, cut, loads, stats = h.uncoarsen(init, og.refine_opts(k))
                runs.append((levels, maps, lab, cut, loads, stats))
            (l1, m1, a1, c1, d1, s1), (l2, m2, a2, c2, d2, s2) = runs
            assert len(l1) == len(l2) and c1 == c2 and s1 == s2
            assert all(all(np.array_equal(x, y) for x, y in zip(p, q)) for p, q in zip(l1, l2))
            assert all(np.array_equal(x, y) for x, y in zip(m1, m2))
            assert np.array_equal(a1, a2) and np.array_equal(d1, d2)
    finally:
        og.set_threads(full)
