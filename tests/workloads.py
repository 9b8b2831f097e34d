"""Mixed-size and served workloads shared by bench.py and the GPU tests.

batch_mix: the batched workload -- `count` graphs of uneven size, 1/3 2-D
grids (side 32-1024), 1/3 R-MAT (scale 10-18, edge factor 8-16, a=0.57
b=0.19 c=0.19), 1/3 random geometric 2-D graphs (2^10-2^20 points, radius
sqrt(8 / (pi n)), mean degree 8); k drawn from {4, 8, 16, 32, 64}, eps 1.03.
A draw whose integer capacity cannot hold the graph is redrawn (same kind),
so the mix always has `count` graphs.

serving_requests: a Poisson stream over a mix -- cold requests with k in
[2, 128] and every tenth request a repartition of a resident graph after 1%
random edge deletions and as many insertions, warm-started from its current
partition (random_delta, apply_delta: the edge changes and their host
restatement).
"""
import numpy as np


def feasible(gp, n, k, imbalance=1.03):
    """floor(eps n / k) must hold ceil(n / k), and the graph must have at
    least two vertices per part (the reference rejects the request
    otherwise)."""
    if n < 2 * k:
        return False
    try:
        gp.capacity(n, k, imbalance)
        return True
    except RuntimeError:
        return False


def draw_graph(gp, kind, rng):
    if kind == 0:
        side = int(rng.integers(32, 1025))
        off, nbr = gp.generate_grid(side, side)
        return "grid%d" % side, off, nbr
    if kind == 1:
        scale, ef = int(rng.integers(10, 19)), int(rng.integers(8, 17))
        off, nbr = gp.generate_rmat(scale, ef, 0.57, 0.19, 0.19, seed=int(rng.integers(1, 1 << 30)))
        return "rmat%d_ef%d" % (scale, ef), off, nbr
    n = int(round(2.0 ** rng.uniform(10, 20)))
    off, nbr = gp.generate_rgg2d(n, (8.0 / (np.pi * n)) ** 0.5, seed=int(rng.integers(1, 1 << 30)))
    return "rgg%d" % n, off, nbr


def batch_mix(gp, count, seed=5):
    """[(name, off, nbr, k)] of exactly `count` feasible graphs."""
    rng = np.random.default_rng(seed)
    mix = []
    for i in range(count):
        k = int(rng.choice([4, 8, 16, 32, 64]))
        while True:
            name, off, nbr = draw_graph(gp, i % 3, rng)
            if feasible(gp, len(off) - 1, k):
                break
        mix.append(("%s_k%d" % (name, k), off, nbr, k))
    return mix


def serving_requests(gp, mix, count, seed=6):
    """[(kind, mix index, payload, k)]: kind "cold" partitions graph j of the
    mix (payload (off, nbr)) at a k drawn from [2, 128] (lowered until
    feasible); kind "warm" (every tenth request) repartitions the resident
    copy of graph j at its mix k after an edge delta (payload (u, v, insert):
    1% random edge deletions and as many insertions), starting from the
    resident graph's current partition."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        j = int(rng.integers(0, len(mix)))
        _, off, nbr, k0 = mix[j]
        if i % 10 == 9:
            reqs.append(("warm", j, random_delta(off, nbr, 0.01, rng), k0))
        else:
            k = int(rng.integers(2, 129))
            while k > 2 and not feasible(gp, len(off) - 1, k):
                k -= 1
            reqs.append(("cold", j, (off, nbr), k))
    return reqs


def poisson_arrivals(count, rate, seed=6):
    return np.cumsum(np.random.default_rng(seed + 1000).exponential(1.0 / rate, count))


def random_delta(off, nbr, frac, rng):
    """(u, v, insert): frac*m/2 deletions of existing undirected edges and as
    many insertions of random vertex pairs (u != v), in a random order."""
    n = len(off) - 1
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(off))
    keep = src < nbr
    a, b = src[keep], nbr[keep].astype(np.int64)
    k = max(1, int(frac * len(a)))
    drop = rng.choice(len(a), size=min(k, len(a)), replace=False)
    du, dv = a[drop], b[drop]
    iu, iv = rng.integers(0, n, k), rng.integers(0, n, k)
    ok = iu != iv
    iu, iv = iu[ok], iv[ok]
    u = np.concatenate([du, iu]).astype(np.int32)
    v = np.concatenate([dv, iv]).astype(np.int32)
    ins = np.concatenate([np.zeros(len(du), np.uint8), np.ones(len(iu), np.uint8)])
    order = rng.permutation(len(u))
    return u[order], v[order], ins[order]


def apply_delta(off, nbr, ew, u, v, insert, w=None):
    """Host restatement of gp_graph_apply_delta: deletions first, then
    insertions; deleting an absent edge changes nothing; inserting an edge
    present after the deletions, or repeated, is dropped (first weight
    kept); each row keeps its surviving entries in their old order followed
    by its new neighbours ascending.  Returns (off, nbr, ew) with ew None
    when every weight is 1 and none was given."""
    n = len(off) - 1
    dels, ins = {}, {}
    for i in range(len(u)):
        a, b = int(u[i]), int(v[i])
        wt = 1 if w is None else int(w[i])
        for s, d in ((a, b), (b, a)):
            if insert[i]:
                ins.setdefault(s, {}).setdefault(d, wt)
            else:
                dels.setdefault(s, set()).add(d)
    weighted = ew is not None or (w is not None and any(int(w[i]) != 1 for i in range(len(u)) if insert[i]))
    new_off = [0]
    new_nbr, new_ew = [], []
    for s in range(n):
        row = nbr[off[s]:off[s + 1]]
        wr = ew[off[s]:off[s + 1]] if ew is not None else None
        gone = dels.get(s, set())
        kept = [(int(x), 1 if wr is None else int(wr[j])) for j, x in enumerate(row) if int(x) not in gone]
        present = {x for x, _ in kept}
        extra = sorted((d, wt) for d, wt in ins.get(s, {}).items() if d not in present)
        for x, wt in kept + extra:
            new_nbr.append(x)
            new_ew.append(wt)
        new_off.append(len(new_nbr))
    return (np.array(new_off, np.int32), np.array(new_nbr, np.int32),
            np.array(new_ew, np.uint32) if weighted else None)
