"""Benchmark: edges partitioned per second on one MI355X.

Workload: R-MAT scale 22 (4,194,304 vertices, edge factor 16,
a=0.57 b=0.19 c=0.19 d=0.05, symmetrised, deduplicated, no self loops, unit
weights), k=64, imbalance 1.03, seed 0.  One step = one full partition
(coarsen -> host initial partition -> uncoarsen) with the input resident in
HBM, i.e. the reference CLI's algorithm_total_seconds scope.  Prints one JSON
line.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

# The CPU baseline's OpenMP threads are pinned to cores; set before the
# oracle library (and its OpenMP runtime) is loaded.
os.environ.setdefault("OMP_PLACES", "cores")
os.environ.setdefault("OMP_PROC_BIND", "close")

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "edges partitioned/s at k=64 ε=3%, batched vs plain; cut ratio; p99 latency"
HBM_PEAK_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--parts", type=int, default=64)
    ap.add_argument("--cpu-scale", type=int, default=22, help="R-MAT scale of the CPU baseline run")
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-graphs", type=int, default=256, help="graphs in the batched mix (0: skip)")
    ap.add_argument("--batch-workers", type=int, default=6)
    ap.add_argument("--serve-requests", type=int, default=128, help="requests of the serving leg (0: skip)")
    ap.add_argument("--serve-detail", default="", help="write the serving leg's per-request records to this JSON file")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three 1 GiB device-to-device copies (counter calibration) and exit")
    ap.add_argument("--trace-mix", action="store_true",
                    help="only run the batched mix through the plain per-graph loop (one host thread, for kernel "
                         "traces of the mixed workload), then exit")
    ap.add_argument("--quality-anchor", action="store_true",
                    help="only partition the 4000x2000 grid at k=4..128 and compare the cuts and times with the "
                         "reference's published run of the same graph, then exit")
    ap.add_argument("--anchor-parts", default="4,8,16,32,64,128", help="parts of the quality anchor")
    ap.add_argument("--anchor-repeats", type=int, default=3)
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.quality_anchor:
        quality_anchor([int(x) for x in args.anchor_parts.split(",")], args.anchor_repeats)
        return
    if args.trace_mix:
        trace_mix(args.batch_graphs)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only; multi-GPU sharding is a design note, not a run.",
              file=sys.stderr)
        sys.exit(2)

    from tests.conftest import load_package
    gp = load_package()

    t = time.time()
    off, nbr = gp.generate_rmat(args.scale, 16, 0.57, 0.19, 0.19, seed=1)
    n, m = len(off) - 1, len(nbr)
    edges = m // 2
    print("bench: rmat scale %d n=%d undirected_edges=%d generated in %.1fs" % (args.scale, n, edges, time.time() - t),
          file=sys.stderr, flush=True)

    ctx = gp.Context(0)
    g = gp.Graph(ctx, off, nbr)
    for i in range(args.warmup):
        _, st = g.partition(args.parts, 1.03, 0, download=False)
        print("bench: warmup %d %.3fs cut=%d" % (i, st["algorithm_total_seconds"], st["final_cut"]),
              file=sys.stderr, flush=True)

    ctx.profile(True)
    ctx.profile_reset()
    per_step = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        s0 = time.perf_counter()
        _, st = g.partition(args.parts, 1.03, 0, download=False)
        per_step.append(time.perf_counter() - s0)
        phases = {"coarsen_ms": 1000.0 * st["coarsen_seconds"],
                  "initial_partition_ms": 1000.0 * st["initial_partition_seconds"],
                  "uncoarsen_ms": 1000.0 * st["uncoarsen_seconds"], "launches": st["launches"],
                  "host_syncs": st["host_syncs"]}
        print("bench: step %d %.3fs (algorithm %.3fs) cut=%d" % (i, per_step[-1], st["algorithm_total_seconds"],
                                                                   st["final_cut"]), file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    prof = {name: ctx.profile_query(name) for name, _, _ in ROOFLINE_CLASSES}
    ctx.profile(False)

    # Verification of one result: exact recount and strict capacity.
    labels, st = g.partition(args.parts, 1.03, 0, download=True)
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(off))
    recount = int(np.count_nonzero(labels[src] != labels[nbr])) // 2
    loads = np.bincount(labels, minlength=args.parts)
    cap = gp.capacity(n, args.parts, 1.03)
    if recount != st["final_cut"] or loads.max() > cap:
        raise RuntimeError("bench: result failed verification")

    value = edges * args.steps / elapsed
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {
            "workload": "rmat-scale%d-ef16-plain-k%d" % (args.scale, args.parts),
            "vertices": n,
            "undirected_edges": edges,
            "parts": args.parts,
            "imbalance": 1.03,
            "seed": 0,
            "cut": int(st["final_cut"]),
            "cut_ratio": st["final_cut"] / edges,
            "max_part_weight": int(st["max_part_weight"]),
            "capacity": int(cap),
            "levels": st["levels"],
            "p50_step_ms": 1000.0 * float(np.percentile(per_step, 50)),
            "p99_step_ms": 1000.0 * float(np.percentile(per_step, 99)),
            "last_step_phases": phases,
        },
    }
    # Rooflines of the row-sweep classes, each timed as one unit (its tier
    # launches) with HIP events on the library's stream; "roofline" is the
    # LP proposal sweep, the class the headline fraction is quoted on.
    for name, key, kernels in ROOFLINE_CLASSES:
        ms, byts, launches = prof[name]
        if not launches:
            continue
        avg_s = ms / 1000.0 / launches
        achieved = byts / launches / avg_s / 1e9
        out[key] = {"bound": "hbm", "kernel": kernels, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_of(name), "avg_launch_us": avg_s * 1e6,
                    "algorithmic_bytes_per_launch": byts / launches, "launches": launches}
    if args.batch_graphs > 0:
        out["batched"], mix, items = batched_leg(gp, ctx, args.batch_graphs, args.batch_workers)
        if args.serve_requests > 0:
            out["serving"] = serving_leg(gp, ctx, mix, items, args.batch_workers, args.serve_requests,
                                         detail=args.serve_detail)
    if not args.no_cpu:
        from oracle import oracle as og
        out["cpu_baseline"] = cpu_baseline(gp, og, args.cpu_scale, args.parts, args.cpu_repeats)
    g.close()
    ctx.close()
    print(json.dumps(out), flush=True)


ROOFLINE_CLASSES = (
    ("lp_propose", "roofline", "k_tier_{light,medium,heavy}<LpProposeBody>"),
    ("contract_scatter", "roofline_contraction", "k_tier_{light,medium,heavy}<ScatterBody>"),
    ("after", "roofline_afterburner", "k_tier_{light,medium,heavy}<AfterBody>"),
    ("update", "roofline_update", "k_tier_{light,medium,heavy}<UpdateBody> + k_commit"),
    ("repair_propose", "roofline_repair", "k_tier_{light,medium,heavy}<RepairProposeBody>"),
)


def trace_mix(count):
    """The mixed workload through the plain per-graph loop on one host
    thread: the form the kernel traces of the mix are taken from."""
    from tests import workloads
    from tests.conftest import load_package
    gp = load_package()
    mix = workloads.batch_mix(gp, count)
    ctx = gp.Context(0)
    t0 = time.perf_counter()
    for i, (name, off, nbr, k) in enumerate(mix):
        g = gp.Graph(ctx, off, nbr)
        g.partition(k)
        g.close()
        if i % 32 == 31:
            print("trace-mix: %d graphs, %.1fs" % (i + 1, time.perf_counter() - t0), file=sys.stderr, flush=True)
    ctx.close()
    print(json.dumps({"trace_mix_graphs": len(mix), "seconds": time.perf_counter() - t0}), flush=True)


def batched_leg(gp, ctx, count, workers):
    """The same mix partitioned by the plain per-graph loop (one context,
    upload + partition + download per graph) and by the batched form
    (concurrent requests); results must be identical."""
    from tests import workloads
    t = time.time()
    mix = workloads.batch_mix(gp, count)
    edges = sum(len(nbr) // 2 for _, _, nbr, _ in mix)
    print("bench: batched mix of %d graphs, %d undirected edges, generated in %.1fs" % (len(mix), edges, time.time() - t),
          file=sys.stderr, flush=True)
    # Warm both paths once on a small graph.
    small = [gp.BatchItem(mix[0][1], mix[0][2], mix[0][3])]
    gp.partition_batch(small, workers=workers)
    plain_labels, plain_cuts, plain_lat = [], [], []
    t0 = time.perf_counter()
    for _, off, nbr, k in mix:
        s0 = time.perf_counter()
        g = gp.Graph(ctx, off, nbr)
        lab, st = g.partition(k)
        g.close()
        plain_lat.append(time.perf_counter() - s0)
        plain_labels.append(lab)
        plain_cuts.append(st["final_cut"])
    plain_s = time.perf_counter() - t0
    items = [gp.BatchItem(off, nbr, k) for _, off, nbr, k in mix]
    t0 = time.perf_counter()
    stats = gp.partition_batch(items, workers=workers)
    batch_s = time.perf_counter() - t0
    for it, st in zip(items, stats):
        it.final_cut = st["final_cut"]
    same = all(np.array_equal(it.labels, lab) for it, lab in zip(items, plain_labels))
    cut_ratio = [(st["final_cut"] + 1) / (c + 1) for st, c in zip(stats, plain_cuts)]
    lat = [st["device_seconds"] for st in stats]
    print("bench: plain %.2fs batched %.2fs identical=%s" % (plain_s, batch_s, same), file=sys.stderr, flush=True)
    if not same:
        raise RuntimeError("bench: batched results differ from the plain loop")
    return {"workload": "mix%d-seed5" % len(mix), "graphs": len(mix), "undirected_edges": edges,
            "seconds": batch_s, "plain_seconds": plain_s,
            "value": edges / batch_s, "plain_value": edges / plain_s, "speedup": plain_s / batch_s,
            "workers": workers, "identical_to_plain": same,
            "cut_geomean_ratio": float(np.exp(np.mean(np.log(cut_ratio)))),
            "p50_request_ms": 1000.0 * float(np.percentile(lat, 50)),
            "p99_request_ms": 1000.0 * float(np.percentile(lat, 99)),
            "plain_p99_request_ms": 1000.0 * float(np.percentile(plain_lat, 99)),
            "launches_per_graph": float(np.mean([st["launches"] for st in stats])),
            "host_syncs_per_graph": float(np.mean([st["host_syncs"] for st in stats])),
            "initial_partition_ms_per_graph": 1000.0 * float(np.mean([st["initial_partition_seconds"]
                                                                       for st in stats]))}, mix, items


def serving_leg(gp, ctx, mix, done_items, workers, count, load=0.7, seed=6, detail=""):
    """Poisson arrivals at `load` times the stream's own capacity (its cold
    requests partitioned back to back by the batched form, in requests/s):
    cold requests from the mix with k in [2, 128], and 10% repartitions of
    resident graphs (kept in HBM by the server with their current
    partition) after a 1% edge delta applied on the device, warm-started
    from that partition.  Latency = queueing + (upload) + partition +
    download.  Quality: each repartition's cut against a cold partition of
    the same updated graph and against the cold cut of the graph before its
    deltas."""
    from tests import workloads
    reqs = workloads.serving_requests(gp, mix, count, seed)
    cold = [gp.BatchItem(p[0], p[1], k) for kind, _, p, k in reqs if kind == "cold"]
    t0 = time.perf_counter()
    gp.partition_batch(cold, workers=workers)
    capacity = len(cold) / (time.perf_counter() - t0)
    rate = load * capacity
    srv = gp.Server(workers=workers)
    try:
        resident = {}
        for kind, j, _, _ in reqs:  # set-up before the stream: resident graphs and their partitions
            if kind == "warm" and j not in resident:
                _, off, nbr, k0 = mix[j]
                resident[j] = srv.add_graph(off, nbr)
                srv.wait(srv.submit(gp.ResidentItem(resident[j], len(off) - 1, k0)))
        items = []
        for kind, j, payload, k in reqs:
            if kind == "warm":
                items.append(gp.ResidentItem(resident[j], len(mix[j][1]) - 1, k, delta=payload + (None,), warm=True))
            else:
                items.append(gp.BatchItem(payload[0], payload[1], k))
        arrivals = workloads.poisson_arrivals(count, rate, seed)
        tickets = []
        t0 = time.perf_counter()
        for item, at in zip(items, arrivals):
            delay = t0 + at - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets.append(srv.submit(item))
        stats = [srv.wait(t) for t in tickets]
        wall = time.perf_counter() - t0
    finally:
        srv.close()
    lat = [st["queue_seconds"] + st["device_seconds"] for st in stats]
    if detail:
        with open(detail, "w") as f:
            json.dump([{"kind": kind, "graph": mix[j][0], "k": k, "arrival_s": float(at), "queue_s": st["queue_seconds"],
                        "device_s": st["device_seconds"], "initial_partition_s": st["initial_partition_seconds"],
                        "launches": st["launches"], "host_syncs": st["host_syncs"]}
                       for (kind, j, _, k), at, st in zip(reqs, arrivals, stats)], f)
    edges = 0
    host = {j: (mix[j][1], mix[j][2], None) for j in resident}
    vs_cold, vs_original = [], []
    for (kind, j, payload, k), st in zip(reqs, stats):
        if kind == "cold":
            edges += len(payload[1]) // 2
            continue
        off, nbr, ew = workloads.apply_delta(*host[j], *payload)
        host[j] = (off, nbr, ew)
        edges += len(nbr) // 2
        g = gp.Graph(ctx, off, nbr, ew)
        _, cold = g.partition(k)
        g.close()
        vs_cold.append((st["final_cut"] + 1) / (cold["final_cut"] + 1))
        vs_original.append((st["final_cut"] + 1) / (done_items[j].final_cut + 1))
    print("bench: served %d requests in %.2fs, p99 %.1f ms" % (count, wall, 1000 * np.percentile(lat, 99)),
          file=sys.stderr, flush=True)
    geo = lambda xs: float(np.exp(np.mean(np.log(xs)))) if xs else None
    mean = lambda key: float(np.mean([st[key] for st in stats]))
    return {"requests": count, "capacity_per_s": capacity, "rate_per_s": rate, "workers": workers,
            "edges_per_s": edges / wall,
            "p50_ms": 1000.0 * float(np.percentile(lat, 50)), "p99_ms": 1000.0 * float(np.percentile(lat, 99)),
            "mean_queue_ms": 1000.0 * mean("queue_seconds"), "mean_device_ms": 1000.0 * mean("device_seconds"),
            "mean_initial_partition_ms": 1000.0 * mean("initial_partition_seconds"),
            "repartitions": len(vs_cold), "resident_graphs": len(resident),
            "warm_vs_cold_cut_geomean": geo(vs_cold), "warm_vs_original_cold_cut_geomean": geo(vs_original)}


# The reference's own run of its partitioner on the 4000x2000 grid
# (15,994,000 edges): k -> (cut, seconds), the "ours" columns of
# results/manual/regret-20261017/comparison.csv:44-49 (median of 3 repeats on
# an RTX 5090), and Jet's cut from the same rows.
GRID_ANCHOR = {4: (8620, 0.068675156, 8228), 8: (14582, 0.071301289, 13798), 16: (24194, 0.090550177, 23186),
               32: (36898, 0.088336934, 34214), 64: (55214, 0.09714246, 51845), 128: (77453, 0.12128948, 73225)}


def quality_anchor(parts, repeats=3):
    """Cut quality against the reference on the one published input that can
    be regenerated here: the 4000x2000 4-neighbour grid.  Vertex numbering
    (row-major here) and the initial partitioner (METIS 5.2.1 there, the
    in-repo bisection here) differ, so the ratio is a sanity anchor."""
    from tests.conftest import load_package
    gp = load_package()
    off, nbr = gp.generate_grid(4000, 2000)
    edges = len(nbr) // 2
    ctx = gp.Context(0)
    g = gp.Graph(ctx, off, nbr)
    rows = {}
    for k in sorted(parts):
        ref_cut, ref_s, jet_cut = GRID_ANCHOR[k]
        g.partition(k, 1.03, 0, download=False)
        secs = []
        for _ in range(repeats):
            _, st = g.partition(k, 1.03, 0, download=False)
            secs.append(st["algorithm_total_seconds"])
        labels, st = g.partition(k, 1.03, 0, download=True)
        src = np.repeat(np.arange(len(off) - 1, dtype=np.int64), np.diff(off))
        recount = int(np.count_nonzero(labels[src] != labels[nbr])) // 2
        cap = gp.capacity(len(off) - 1, k, 1.03)
        if recount != st["final_cut"] or np.bincount(labels, minlength=k).max() > cap:
            raise RuntimeError("quality anchor: result failed verification")
        rows[k] = {"cut": int(st["final_cut"]), "reference_cut": ref_cut, "cut_ratio_vs_reference": st["final_cut"] / ref_cut,
                   "jet_cut": jet_cut, "seconds_median": float(np.median(secs)), "reference_seconds": ref_s,
                   "speedup_vs_reference": ref_s / float(np.median(secs)), "levels": st["levels"]}
        print("anchor: grid 4000x2000 k=%d cut=%d (reference %d, ratio %.4f) %.1f ms (reference %.1f ms)" % (
            k, st["final_cut"], ref_cut, st["final_cut"] / ref_cut, 1e3 * np.median(secs), 1e3 * ref_s),
              file=sys.stderr, flush=True)
    g.close()
    ctx.close()
    geo = float(np.exp(np.mean([np.log(r["cut_ratio_vs_reference"]) for r in rows.values()])))
    print(json.dumps({"anchor": "grid-4000x2000", "undirected_edges": edges, "imbalance": 1.03, "seed": 0,
                      "cut_geomean_ratio_vs_reference": geo, "parts": rows}), flush=True)


def copy_calibration():
    """Three plain 1 GiB device copies: what FETCH_SIZE / WRITE_SIZE read
    for a known byte count scales the counters of the partitioner."""
    import torch
    a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copy calibration: 3 x %d bytes" % (4 << 28), flush=True)


def traffic_of(name):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_%s_rmat22_k64.txt" % name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch="):
            return float(line.split("=", 1)[1])
    return None


def cpu_partition(gp, og, off, nbr, parts):
    """The CPU path: the oracle's coarsening (OpenMP), the host initial
    partition of the coarsest graph, the oracle's uncoarsening; seconds in
    the algorithm_total_seconds scope (graph in memory -> labels in memory)."""
    t = time.perf_counter()
    oh = og.Hierarchy(off, nbr, None, None, og.coarsen_opts(parts))
    coff, cnbr, cew, cvw = oh.graph(oh.levels - 1)
    init = gp.initial_partition_host(coff, cnbr, cew, cvw, parts)
    labels, cut, loads, _ = oh.uncoarsen(init, og.refine_opts(parts))
    return time.perf_counter() - t, labels, cut


def verify_with_checker(gp, off, nbr, labels, parts, cut):
    """gpart_verify_csr on the written CSR and partition: exact cut and loads
    within the integer capacity."""
    tool = os.path.join(ROOT, "acc-gp_amd", "bin", "gpart_verify_csr")
    with tempfile.TemporaryDirectory() as d:
        ip, ix, pp = os.path.join(d, "g.indptr"), os.path.join(d, "g.indices"), os.path.join(d, "part.bin")
        gp.save_csr(ip, ix, off, nbr)
        np.asarray(labels, "<i4").tofile(pp)
        res = subprocess.run([tool, ip, ix, str(parts), pp, "binary", "--symmetric"], capture_output=True, text=True)
    got = json.loads(res.stdout)
    ok = res.returncode == 0 and got["cut"] == cut and max(got["loads"]) <= got["capacity"]
    if not ok:
        raise RuntimeError("cpu baseline: gpart_verify_csr rejected the result: %s" % res.stdout)
    return got


def cpu_baseline(gp, og, scale, parts, repeats):
    """The repo's CPU path on the box's host cores: the OpenMP oracle
    (the same schedule as the device) with the host initial partitioner, on
    the 256x256 grid at k=8 and on the bench's R-MAT configuration; median
    of `repeats` runs (one when a run exceeds 20 s), every result checked by
    gpart_verify_csr."""
    threads = og.get_threads()
    out = {"unit": "edges/s", "cores": threads, "kind": "port", "nproc": os.cpu_count(),
           "omp_places": os.environ.get("OMP_PLACES"), "omp_proc_bind": os.environ.get("OMP_PROC_BIND")}
    runs = {}
    for name, (off, nbr), k in (("grid256_k8", gp.generate_grid(256, 256), 8),
                               ("rmat%d_k%d" % (scale, parts), gp.generate_rmat(scale, 16, 0.57, 0.19, 0.19, seed=1),
                                parts)):
        secs, cut = [], None
        for r in range(repeats):
            s, labels, cut = cpu_partition(gp, og, off, nbr, k)
            secs.append(s)
            if r == 0:
                verify_with_checker(gp, off, nbr, labels, k, cut)
            if s > 20.0:
                break
        edges = len(nbr) // 2
        med = float(np.median(secs))
        runs[name] = {"undirected_edges": edges, "parts": k, "seconds_median": med, "repeats": len(secs),
                      "edges_per_s": edges / med, "cut": int(cut), "verified": True}
        print("bench: cpu baseline %s %.2fs (%d threads) cut=%d" % (name, med, threads, cut), file=sys.stderr,
              flush=True)
    main_run = runs["rmat%d_k%d" % (scale, parts)]
    out["value"] = main_run["edges_per_s"]
    out["sample"] = ("R-MAT scale %d ef16 (%d undirected edges) k=%d, median of %d: OpenMP oracle coarsen + host "
                     "initial partition + oracle uncoarsen on %d threads, checked by gpart_verify_csr; also the "
                     "256x256 grid at k=8" % (scale, main_run["undirected_edges"], parts, main_run["repeats"],
                                              threads))
    out["runs"] = runs
    return out


if __name__ == "__main__":
    main()
