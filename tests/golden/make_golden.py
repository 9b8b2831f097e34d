"""Writes the full-size golden fixtures: the oracle's run of the schedule on
the full-size workloads, stored as SHA-256 digests.

    python tests/golden/make_golden.py [name ...]

Each fixture holds, for one workload (R-MAT scale 22 at k=64; the weighted
3-D 10-nearest-neighbour graph of 8M points at k=128): the digest of the
generated input CSR, per coarsening level the vertex and entry counts, the
level metadata and the digests of the offsets, neighbours, edge weights,
vertex weights and the fine-to-coarse map; the digest of the host initial
partition of the coarsest graph; and the digest of the refined labels, the
cut, the part loads and every per-level refinement statistic.  The oracle
takes minutes at these sizes, so the fixtures are written once here and the
GPU test compares the device run against them (tests/test_gpu_golden.py).
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as og  # noqa: E402
from tests.conftest import load_package  # noqa: E402


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        if a is None:
            h.update(b"none")
        else:
            h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


WORKLOADS = {
    "rmat22_k64": dict(kind="rmat", scale=22, edge_factor=16, seed=1, parts=64, imbalance=1.03),
    "knn8m_k128": dict(kind="knn3d", points=1 << 23, knn=10, seed=2, parts=128, imbalance=1.03),
}


def generate(gp, w):
    if w["kind"] == "rmat":
        off, nbr = gp.generate_rmat(w["scale"], w["edge_factor"], seed=w["seed"])
        return off, nbr, None, None
    return gp.generate_knn3d(w["points"], w["knn"], seed=w["seed"])


def make(name):
    gp = load_package()
    w = WORKLOADS[name]
    t0 = time.time()
    off, nbr, ew, vw = generate(gp, w)
    k = w["parts"]
    out = dict(workload=name, spec=w, n=len(off) - 1, m=len(nbr), input=digest(off, nbr, ew, vw), levels=[])
    print("%s: generated n=%d m=%d (%.0f s)" % (name, out["n"], out["m"], time.time() - t0), flush=True)
    h = og.Hierarchy(off, nbr, ew, vw, og.coarsen_opts(k))
    print("%s: oracle coarsened to %d levels (%.0f s)" % (name, h.levels, time.time() - t0), flush=True)
    for lvl in range(h.levels):
        info = h.info(lvl)
        goff, gnbr, gew, gvw = h.graph(lvl)
        entry = dict(n=info["n"], m=info["m"], max_degree=info["max_degree"], has_heavy=info["has_heavy"],
                     incident_bound=info["incident_bound"], offsets=digest(goff), neighbors=digest(gnbr),
                     edge_weights=digest(gew), vertex_weights=digest(gvw))
        if lvl + 1 < h.levels:
            entry["map"] = digest(h.map(lvl))
        out["levels"].append(entry)
    coff, cnbr, cew, cvw = h.graph(h.levels - 1)
    init = gp.initial_partition_host(coff, cnbr, cew, cvw, k)
    out["initial_partition"] = digest(init)
    lab, cut, loads, stats = h.uncoarsen(init, og.refine_opts(k, w["imbalance"]))
    print("%s: oracle refined, cut %d (%.0f s)" % (name, cut, time.time() - t0), flush=True)
    out["labels"] = digest(lab)
    out["cut"] = int(cut)
    out["loads"] = [int(x) for x in loads]
    out["level_stats"] = [{key: int(v) for key, v in s.items()} for s in stats]
    path = os.path.join(HERE, name + ".json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")
    print("%s: wrote %s (%.0f s)" % (name, path, time.time() - t0), flush=True)


if __name__ == "__main__":
    for name in sys.argv[1:] or sorted(WORKLOADS):
        make(name)
