"""Host-side checks that need no GPU: the C ABI library and its exports, the
graph generators and the on-disk CSR format, the host initial partitioner,
and the weighted CSR checker against the reference's own gpart_verify_csr
(tools/verify_csr.cpp, built into oracle/_ref when the reference tree is
present)."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

from tests import graphs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gpart_mi355x.h")
VERIFY = os.path.join(ROOT, "acc-gp_amd", "bin", "gpart_verify_csr")
REF_VERIFY = os.path.join(ROOT, "oracle", "_ref", "gpart_verify_csr_ref")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(gp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(gp):
    lib = gp.lib()
    names = _declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(gp.EXPORTS) == sorted(set(gp.EXPORTS))
    assert set(gp.EXPORTS) <= set(names)


def test_library_symbols_by_nm():
    so = os.path.join(ROOT, "acc-gp_amd", "libgpart_mi355x.so")
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (gp_[a-z0-9_]+)", out))
    assert set(_declared()) <= exported


def test_capacity_through_the_abi(gp, og):
    for total, parts, eps in [(65536, 8, 1.03), (4194304, 64, 1.03), (20, 2, 1.0), (12345678901, 64, 1.03)]:
        assert gp.capacity(total, parts, eps) == og.capacity(total, parts, eps)
    with pytest.raises(Exception):
        gp.capacity(7, 3, 1.03)
    with pytest.raises(Exception):
        gp.capacity(100, 1, 1.03)


def _check_sym_csr(off, nbr):
    n = len(off) - 1
    assert off[0] == 0 and off[-1] == len(nbr) and np.all(np.diff(off) >= 0)
    src = np.repeat(np.arange(n), np.diff(off))
    assert not np.any(src == nbr)
    for v in range(0, n, max(1, n // 200)):
        row = nbr[off[v]:off[v + 1]]
        assert np.all(np.diff(row) > 0)
    a = set(zip(src.tolist(), nbr.tolist()))
    assert all((y, x) in a for x, y in a)


def test_generate_grid_matches_reference_numbering(gp):
    off, nbr = gp.generate_grid(256, 256)
    assert len(off) - 1 == 65536 and len(nbr) == 2 * 130560
    ref_off, ref_nbr = graphs.grid(256, 256)
    assert np.array_equal(off, ref_off) and np.array_equal(nbr, ref_nbr)


def test_generate_rmat_properties(gp):
    off, nbr = gp.generate_rmat(12, 16, seed=1)
    assert len(off) - 1 == 4096
    _check_sym_csr(off, nbr)
    again = gp.generate_rmat(12, 16, seed=1)
    assert np.array_equal(off, again[0]) and np.array_equal(nbr, again[1])
    other = gp.generate_rmat(12, 16, seed=2)
    assert not np.array_equal(nbr[:1000], other[1][:1000])
    deg = np.diff(off)
    assert deg.max() > 20 * max(1, int(np.median(deg))), "skewed degrees"


def test_generate_knn_and_rgg(gp):
    off, nbr, ew, vw = gp.generate_knn3d(2000, 6, seed=2)
    _check_sym_csr(off, nbr)
    assert ew.min() >= 1 and ew.max() <= 100 and vw.min() >= 1 and vw.max() <= 4
    src = np.repeat(np.arange(2000), np.diff(off))
    w = dict(zip(zip(src.tolist(), nbr.tolist()), ew.tolist()))
    assert all(w[(b, a)] == x for (a, b), x in w.items())
    off, nbr = gp.generate_rgg2d(4096, (8.0 / (3.14159265 * 4096)) ** 0.5, seed=5)
    _check_sym_csr(off, nbr)


def test_csr_file_roundtrip(gp, tmp_path):
    off, nbr = gp.generate_rmat(10, 8, seed=3)
    ip, ix = str(tmp_path / "g.indptr"), str(tmp_path / "g.indices")
    gp.save_csr(ip, ix, off, nbr)
    assert os.path.getsize(ip) == 8 * len(off) and os.path.getsize(ix) == 8 * len(nbr)
    assert np.array_equal(np.fromfile(ip, "<i8"), off) and np.array_equal(np.fromfile(ix, "<i8"), nbr)
    o2, n2 = gp.load_csr(ip, ix)
    assert np.array_equal(o2, off) and np.array_equal(n2, nbr)


def test_csr_loader_rejects_bad_input(gp, tmp_path):
    ip, ix = str(tmp_path / "b.indptr"), str(tmp_path / "b.indices")
    np.array([0, 2, 3], "<i8").tofile(ip)
    np.array([1, 5, 0], "<i8").tofile(ix)  # neighbour id out of range
    with pytest.raises(Exception):
        gp.load_csr(ip, ix)
    np.array([0, 3, 2], "<i8").tofile(ip)  # not monotone
    np.array([1, 0, 0], "<i8").tofile(ix)
    with pytest.raises(Exception):
        gp.load_csr(ip, ix)


def test_initial_partition_host_valid_and_deterministic(gp, og):
    for k in (2, 8, 64):
        off, nbr = gp.generate_grid(24, 24)
        n = len(off) - 1
        a = gp.initial_partition_host(off, nbr, None, None, k)
        b = gp.initial_partition_host(off, nbr, None, None, k)
        assert np.array_equal(a, b)
        assert a.min() >= 0 and a.max() < k
    off, nbr, ew = graphs.random_graph(500, 2000, seed=3, max_weight=5)
    vw = np.random.default_rng(1).integers(1, 5, 500).astype(np.uint32)
    lab = gp.initial_partition_host(off, nbr, ew, vw, 8)
    assert lab.min() >= 0 and lab.max() < 8


def _write_partition(tmp_path, labels, fmt):
    p = str(tmp_path / ("part." + fmt))
    if fmt == "binary":
        np.asarray(labels, "<i4").tofile(p)
    else:
        open(p, "w").write("\n".join(str(int(x)) for x in labels) + "\n")
    return p


@pytest.mark.parametrize("fmt", ["binary", "text"])
def test_verify_tool_matches_reference_checker(gp, og, tmp_path, fmt):
    off, nbr = gp.generate_rmat(11, 8, seed=6)
    n, k = len(off) - 1, 8
    ip, ix = str(tmp_path / "g.indptr"), str(tmp_path / "g.indices")
    gp.save_csr(ip, ix, off, nbr)
    lab = (np.arange(n) % k).astype(np.int32)
    part = _write_partition(tmp_path, lab, fmt)
    res = subprocess.run([VERIFY, ip, ix, str(k), part, fmt, "--symmetric"], capture_output=True, text=True)
    got = json.loads(res.stdout)
    assert got["cut"] == og.cut(off, nbr, None, lab)
    assert got["undirected_edges"] == len(nbr) // 2
    assert got["loads"] == np.bincount(lab, minlength=k).tolist()
    assert got["capacity"] == n * 103 // (100 * k)
    if os.path.exists(REF_VERIFY):
        ref = subprocess.run([REF_VERIFY, ip, ix, str(k), part, fmt, "--symmetric"], capture_output=True, text=True)
        want = json.loads(ref.stdout)
        assert res.returncode == ref.returncode
        for key, value in want.items():
            assert got[key] == value, key


def test_verify_tool_weighted_extension(gp, og, tmp_path):
    off, nbr, ew, vw = gp.generate_knn3d(1500, 6, seed=2)
    n, k = len(off) - 1, 4
    ip, ix = str(tmp_path / "w.indptr"), str(tmp_path / "w.indices")
    gp.save_csr(ip, ix, off, nbr)
    ewf, vwf = str(tmp_path / "w.ew"), str(tmp_path / "w.vw")
    ew.astype("<u4").tofile(ewf)
    vw.astype("<u4").tofile(vwf)
    lab = (np.arange(n) % k).astype(np.int32)
    part = _write_partition(tmp_path, lab, "binary")
    res = subprocess.run([VERIFY, ip, ix, str(k), part, "binary", "--edge-weights", ewf, "--vertex-weights", vwf],
                         capture_output=True, text=True)
    got = json.loads(res.stdout)
    assert got["cut"] == og.cut(off, nbr, ew, lab)
    assert got["loads"] == np.bincount(lab, weights=vw, minlength=k).astype(np.int64).tolist()


def test_server_fails_cleanly_without_a_device(gp):
    # Failure handling runs on the CPU: with no usable device every worker
    # fails to create its context and requests report an error (no hang).
    try:
        gp.Context(0).close()
        pytest.skip("a GPU is present")
    except RuntimeError:
        pass
    off, nbr = gp.generate_grid(8, 8)
    with pytest.raises(RuntimeError):
        gp.partition_batch([gp.BatchItem(off, nbr, 2)], workers=2)


# ------------------------------------------------ initial-partition boundary
def _repair_brute(off, nbr, ew, vw, k, cap, lab):
    """repair_capacity of src/initial_partition_metis.cu:35-83 in exact
    rationals: while a part is above capacity, move the (vertex, part) of
    largest (A(v,p) - A(v,s)) / min(w, load_s - cap), first in (v, p) order."""
    from fractions import Fraction
    n = len(off) - 1
    lab = list(lab)
    w = [1] * n if vw is None else [int(x) for x in vw]
    loads = [0] * k
    for v in range(n):
        loads[lab[v]] += w[v]
    moved = 0
    while max(loads) > cap:
        best, pick = None, None
        for v in range(n):
            s = lab[v]
            if loads[s] <= cap:
                continue
            a = [0] * k
            for e in range(off[v], off[v + 1]):
                if nbr[e] != v:
                    a[lab[nbr[e]]] += 1 if ew is None else int(ew[e])
            for p in range(k):
                if p == s or loads[p] > cap or w[v] > cap - loads[p]:
                    continue
                score = Fraction(a[p] - a[s], min(w[v], loads[s] - cap))
                if best is None or score > best:
                    best, pick = score, (v, p)
        if pick is None:
            raise RuntimeError("no feasible move")
        v, p = pick
        loads[lab[v]] -= w[v]
        loads[p] += w[v]
        lab[v] = p
        moved += 1
    return lab, moved


def test_repair_capacity_known_answer(gp):
    # Path 0-1-2-3, part 0 = {0,1,2} over capacity 2: vertex 2 (gain 0) beats
    # vertex 0 (gain -1) and vertex 1 (gain -2).
    off, nbr, _ = graphs.csr_from_edges(4, [(0, 1), (1, 2), (2, 3)])
    lab, moved = gp.repair_capacity_host(off, nbr, None, None, 2, 2, [0, 0, 0, 1])
    assert lab.tolist() == [0, 0, 1, 1] and moved == 1
    # Already feasible: nothing moves.
    lab, moved = gp.repair_capacity_host(off, nbr, None, None, 2, 2, [0, 0, 1, 1])
    assert lab.tolist() == [0, 0, 1, 1] and moved == 0
    # No part has room: the strict repair fails.
    with pytest.raises(RuntimeError, match="strict capacity"):
        gp.repair_capacity_host(off, nbr, None, np.array([3, 3, 3, 3], np.uint32), 2, 4, [0, 0, 0, 1])


@pytest.mark.parametrize("seed", range(6))
def test_repair_capacity_brute_force(gp, seed):
    rng = np.random.default_rng(seed)
    n, k = 40, 4
    off, nbr, ew = graphs.random_graph(n, 90, seed=seed + 100, max_weight=5)
    vw = rng.integers(1, 5, n).astype(np.uint32)
    lab = rng.integers(0, k, n).astype(np.int32)
    lab[rng.random(n) < 0.4] = 1
    cap = int(np.ceil(vw.sum() * 1.1 / k))
    got, moved = gp.repair_capacity_host(off, nbr, ew, vw, k, cap, lab)
    want, wmoved = _repair_brute(off, nbr, ew, vw, k, cap, lab)
    assert got.tolist() == want and moved == wmoved
    assert np.bincount(got, weights=vw, minlength=k).max() <= cap


def test_initial_partition_gcd_normalisation(gp):
    # Common factors of vertex and edge weights are divided out exactly, so
    # scaling every vertex weight by 3 and every edge weight by 5 leaves the
    # labels unchanged; self loops and zero-weight edges are left out.
    off, nbr, ew = graphs.random_graph(500, 2000, seed=9, max_weight=4)
    vw = np.random.default_rng(2).integers(1, 5, 500).astype(np.uint32)
    base = gp.initial_partition_host(off, nbr, ew, vw, 8)
    scaled = gp.initial_partition_host(off, nbr, ew * 5, vw * 3, 8)
    assert np.array_equal(base, scaled)
    # add a self loop on every vertex and a zero-weight edge
    rows = [[(int(nbr[e]), int(ew[e])) for e in range(off[v], off[v + 1])] + [(v, 7)] for v in range(500)]
    rows[0].append((499, 0))
    rows[499].append((0, 0))
    off2, nbr2, ew2 = [0], [], []
    for r in rows:
        for u, w in sorted(r):
            nbr2.append(u)
            ew2.append(w)
        off2.append(len(nbr2))
    looped = gp.initial_partition_host(np.array(off2, np.int32), np.array(nbr2, np.int32),
                                       np.array(ew2, np.uint32), vw, 8)
    assert np.array_equal(base, looped)


def test_initial_partition_boundary_checks(gp):
    # total 130, capacity floor(1.03 * 130 / 2) = 66 < the vertex of weight 70
    off, nbr, _ = graphs.csr_from_edges(61, [(i, i + 1) for i in range(60)])
    with pytest.raises(RuntimeError, match="exceeds partition capacity"):
        gp.initial_partition_host(off, nbr, None, np.array([1] * 30 + [70] + [1] * 30, np.uint32), 2)
    off, nbr, _ = graphs.csr_from_edges(5, [(0, 1), (1, 2), (2, 3), (3, 4)])
    with pytest.raises(RuntimeError, match="invalid initial partition options"):
        gp.initial_partition_host(off, nbr, None, None, 6)
    # weights all 2: capacity floor(10 / 2) = 5 holds 2 vertices (4), below
    # ceil(10 / 2) = 5, so a strict request fails; production does not check.
    even = np.full(5, 2, np.uint32)
    with pytest.raises(RuntimeError, match="indivisible"):
        gp.initial_partition_host(off, nbr, None, even, 2, imbalance=1.0, require_feasible=True)
    lab = gp.initial_partition_host(off, nbr, None, even, 2, imbalance=1.0)
    assert set(lab.tolist()) <= {0, 1}


def test_initial_partition_strict_is_feasible(gp):
    for seed in range(4):
        off, nbr, ew = graphs.random_graph(300, 900, seed=seed, max_weight=3)
        vw = np.random.default_rng(seed).integers(1, 9, 300).astype(np.uint32)
        for k in (2, 5, 16):
            lab = gp.initial_partition_host(off, nbr, ew, vw, k, require_feasible=True)
            cap = gp.capacity(int(vw.sum()), k)
            assert np.bincount(lab, weights=vw, minlength=k).max() <= cap, (seed, k)
    assert gp.initial_partition_backend() in ("metis", "bisection")


def test_edge_delta_host_semantics():
    # The host restatement of gp_graph_apply_delta against the edge-set
    # definition: the new edge set is (E - D) + I, rows keep their surviving
    # entries in order and then the new neighbours ascending, weights of old
    # edges stay, inserted ones carry their first weight.
    from tests import graphs, workloads
    rng = np.random.default_rng(3)
    off, nbr, ew = graphs.random_graph(300, 1200, seed=5, max_weight=9)
    u, v, ins = workloads.random_delta(off, nbr, 0.05, rng)
    # edge cases: delete an absent edge, insert a present one, insert twice, delete and re-insert
    present = (0, int(nbr[off[0]]))
    u = np.concatenate([u, [0, present[0], 7, 7, present[0], present[0]]]).astype(np.int32)
    v = np.concatenate([v, [299 if 299 not in set(nbr[off[0]:off[1]]) else 298, present[1], 9, 9,
                            present[1], present[1]]]).astype(np.int32)
    ins = np.concatenate([ins, [0, 1, 1, 1, 0, 1]]).astype(np.uint8)
    w = rng.integers(1, 50, len(u)).astype(np.uint32)
    off2, nbr2, ew2 = workloads.apply_delta(off, nbr, ew, u, v, ins, w)
    old = {}
    for s in range(len(off) - 1):
        for j in range(off[s], off[s + 1]):
            old[(s, int(nbr[j]))] = int(ew[j])
    dels = {(int(a), int(b)) for a, b, i in zip(u, v, ins) if not i} | {(int(b), int(a)) for a, b, i in zip(u, v, ins) if not i}
    want = {e: x for e, x in old.items() if e not in dels}
    for a, b, i, x in zip(u, v, ins, w):
        if i:
            for e in ((int(a), int(b)), (int(b), int(a))):
                if e not in want:
                    want[e] = int(x)
    got = {}
    for s in range(len(off2) - 1):
        row = [int(x) for x in nbr2[off2[s]:off2[s + 1]]]
        assert len(set(row)) == len(row)
        old_part = [x for x in row if (s, x) in old and (s, x) not in dels]
        assert row[:len(old_part)] == [int(x) for x in nbr[off[s]:off[s + 1]] if (s, int(x)) not in dels]
        assert row[len(old_part):] == sorted(row[len(old_part):])
        for j in range(off2[s], off2[s + 1]):
            got[(s, int(nbr2[j]))] = int(ew2[j])
    assert got == want
    assert got[present] == want[present]  # deleted then re-inserted: the new weight


def test_bench_modes_are_wired():
    # Every mode bench.py dispatches to exists (its GPU legs run on the box).
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    for name in ("main", "quality_anchor", "trace_mix", "copy_calibration", "batched_leg", "serving_leg",
                 "cpu_baseline", "cpu_partition", "verify_with_checker", "traffic_of"):
        assert callable(getattr(bench, name)), name
    assert set(bench.GRID_ANCHOR) == {4, 8, 16, 32, 64, 128}
    assert [c[0] for c in bench.ROOFLINE_CLASSES][0] == "lp_propose"


def test_device_stale_rule_equals_long_double(gp, og):
    """The device loop control decides the small-level stale rule
    cut < 0.999 * best (incremental.cuh:911-961, long double on the host) in
    integers; it must agree with the x87 evaluation everywhere, including
    where the rounded product meets an integer."""
    rng = np.random.default_rng(11)
    cases = []
    for bits in range(1, 64):
        for _ in range(40):
            best = int(rng.integers(1 << (bits - 1), 1 << bits, dtype=np.uint64))
            base = 999 * best // 1000
            cases += [(max(base + d, 0), best) for d in (-2, -1, 0, 1, 2)]
    for best in list(range(1000, 200001, 1000)) + [10 ** 12, 10 ** 15, 10 ** 18, (1 << 63) + 1000, (1 << 64) - 1]:
        base = 999 * best // 1000
        cases += [(max(base + d, 0), best) for d in (-1, 0, 1)]
    cases += [(0, 0), (0, 1), (1, 1), ((1 << 64) - 1, (1 << 64) - 1), (0, (1 << 64) - 1)]
    rational_differs = 0
    for cut, best in cases:
        want = og.below_0999(cut, best)
        assert gp.below_0999(cut, best) == want, (cut, best)
        rational_differs += want != (1000 * cut < 999 * best)
    # The long double constant is not 999/1000: the exact rational rule
    # decides some of these cases differently, so the integer form above is
    # not the rational one.
    assert rational_differs > 0
