"""Small deterministic test graphs (numpy), independent of the library."""
import numpy as np


def csr_from_edges(n, edges, weights=None):
    """Symmetric CSR with sorted rows from undirected (u, v[, w]) edges."""
    rows = [[] for _ in range(n)]
    for i, (u, v) in enumerate(edges):
        w = 1 if weights is None else weights[i]
        rows[u].append((v, w))
        rows[v].append((u, w))
    off = [0]
    nbr, ew = [], []
    for r in rows:
        r.sort()
        for v, w in r:
            nbr.append(v)
            ew.append(w)
        off.append(len(nbr))
    return np.array(off, np.int32), np.array(nbr, np.int32), np.array(ew, np.uint32)


def grid(rows, cols):
    edges = []
    for r in range(rows):
        for c in range(cols):
            v = r * cols + c
            if c + 1 < cols:
                edges.append((v, v + 1))
            if r + 1 < rows:
                edges.append((v, v + cols))
    off, nbr, _ = csr_from_edges(rows * cols, edges)
    return off, nbr


def random_graph(n, m, seed, max_weight=1):
    rng = np.random.default_rng(seed)
    seen = set()
    edges, weights = [], []
    while len(edges) < m:
        u, v = (int(x) for x in rng.integers(0, n, 2))
        if u == v or (min(u, v), max(u, v)) in seen:
            continue
        seen.add((min(u, v), max(u, v)))
        edges.append((u, v))
        weights.append(int(rng.integers(1, max_weight + 1)))
    return csr_from_edges(n, edges, weights)


def hub_ring(n, extra, seed):
    """Vertex 0 joined to every other vertex, a ring over 1..n-1 and `extra`
    random chords: one row far above a block's share (n - 1 entries) beside
    short rows.  Unit edge weights; vertex weights skewed (1..1000,
    log-uniform) so capacity repair has work."""
    rng = np.random.default_rng(seed)
    edges = set()
    for v in range(1, n):
        edges.add((0, v))
        w = v + 1 if v + 1 < n else 1
        edges.add((min(v, w), max(v, w)))
    while len(edges) < (n - 1) * 2 + extra:
        u, v = (int(x) for x in rng.integers(1, n, 2))
        if u != v:
            edges.add((min(u, v), max(u, v)))
    off, nbr, ew = csr_from_edges(n, sorted(edges))
    vw = np.exp(rng.uniform(0.0, np.log(1000.0), n)).astype(np.uint32).clip(1, 1000)
    return off, nbr, ew, vw


def heavy_weights(n, m, seed):
    """Random graph whose edge weights reach 2^30: about 30% of edges weigh
    2^30 + 1000 h (h < 300), the rest 1 + h, with h a hash of the edge.  Rows
    with several heavy edges have incident weights above 2^32, and merged
    rows grow further after contraction.  Vertex weights 1..4."""
    off, nbr, _ = random_graph(n, m, seed=seed)
    src = np.repeat(np.arange(n), np.diff(off))
    lo, hi = np.minimum(src, nbr).astype(np.uint64), np.maximum(src, nbr).astype(np.uint64)
    h = (lo * np.uint64(2654435761) + hi * np.uint64(40503)) % np.uint64(1000)
    ew = np.where(h < 300, np.uint64(1 << 30) + h * np.uint64(1000), np.uint64(1) + h).astype(np.uint32)
    vw = np.random.default_rng(seed + 1).integers(1, 5, n).astype(np.uint32)
    return off, nbr, ew, vw
