"""Mixed-size and served workloads shared by bench.py and the GPU tests.

batch_mix: the batched workload -- `count` graphs of uneven size, 1/3 2-D
grids (side 32-1024), 1/3 R-MAT (scale 10-18, edge factor 8-16, a=0.57
b=0.19 c=0.19), 1/3 random geometric 2-D graphs (2^10-2^20 points, radius
sqrt(8 / (pi n)), mean degree 8); k drawn from {4, 8, 16, 32, 64}, eps 1.03.
A draw whose integer capacity cannot hold the graph is redrawn (same kind),
so the mix always has `count` graphs.

serving_requests: a Poisson stream over a mix -- cold requests with k in
[2, 128] and every tenth request a repartition of an earlier graph after 1%
random edge deletions and as many insertions, warm-started from that
graph's labels.
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


def perturb(off, nbr, frac, rng):
    """Applies frac*m/2 random edge deletions and as many insertions
    (symmetric, sorted rows, no self loops, no duplicates)."""
    n = len(off) - 1
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(off))
    keep = src < nbr
    a, b = src[keep], nbr[keep].astype(np.int64)
    k = max(1, int(frac * len(a)))
    drop = rng.choice(len(a), size=k, replace=False)
    mask = np.ones(len(a), bool)
    mask[drop] = False
    a, b = a[mask], b[mask]
    u, v = rng.integers(0, n, k), rng.integers(0, n, k)
    ok = u != v
    a = np.concatenate([a, np.minimum(u, v)[ok]])
    b = np.concatenate([b, np.maximum(u, v)[ok]])
    key = np.unique(a * n + b)
    a, b = key // n, key % n
    s2 = np.concatenate([a, b])
    d2 = np.concatenate([b, a])
    order = np.lexsort((d2, s2))
    s2, d2 = s2[order], d2[order]
    off2 = np.zeros(n + 1, np.int64)
    np.add.at(off2, s2 + 1, 1)
    return np.cumsum(off2).astype(np.int32), d2.astype(np.int32)


def serving_requests(gp, mix, count, seed=6):
    """[(kind, mix index, off, nbr, k)]: kind "cold" partitions graph j of the
    mix at a k drawn from [2, 128] (lowered until feasible); kind "warm"
    (every tenth request) repartitions a perturbed copy of graph j at its mix
    k from graph j's labels."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        j = int(rng.integers(0, len(mix)))
        _, off, nbr, k0 = mix[j]
        if i % 10 == 9:
            off2, nbr2 = perturb(off, nbr, 0.01, rng)
            reqs.append(("warm", j, off2, nbr2, k0))
        else:
            k = int(rng.integers(2, 129))
            while k > 2 and not feasible(gp, len(off) - 1, k):
                k -= 1
            reqs.append(("cold", j, off, nbr, k))
    return reqs


def poisson_arrivals(count, rate, seed=6):
    return np.cumsum(np.random.default_rng(seed + 1000).exponential(1.0 / rate, count))
