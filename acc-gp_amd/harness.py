"""Run and analysis harness for partitioner binaries: the roles of the
reference's run.sh (run loop, run.sh:173-282) and analyze_cuts.sh (checks and
CSV tables, analyze_cuts.sh:116-182), with the same CSV columns so tables
compare like for like.

    python acc-gp_amd/harness.py run OUT --graph grid:4000x2000 --graph rmat:20:16 \\
        --parts 8,64 --repeats 3 [--method NAME=BINARY ...] [--timeout S]
    python acc-gp_amd/harness.py analyze OUT

run: writes each graph in the reference's int64 CSR format under OUT/graphs
(generated: grid:RxC, rmat:SCALE:EDGEFACTOR[:SEED], rgg:N[:SEED]; or
file:INDPTR:INDICES), then for every graph, k, repeat and method runs the
binary as the reference's CLI is run (<indptr> <indices> <k> <labels.out>
1.03 0), rotating the method order per repeat as run.sh does, and appends one
record per run to OUT/runs.jsonl: status, algorithm_total_seconds and the
coarsen/initial-partition/uncoarsen stage seconds parsed from the
partition_timing line, and the reported final_edge_cut.  The method "ours"
is acc-gp_amd/bin/gpart_partition_standard unless given; any other binary
that takes the same arguments and prints the same lines (the reference's
CLI on its own hardware, an archived build) can be added as NAME=BINARY.

analyze: checks every partition file with gpart_verify_csr (exact cut,
loads, the integer capacity of eps 1.03) and writes OUT/runs.csv,
OUT/comparison.csv (medians over verified repeats per graph, k and method,
and ratios of "ours" over "reference" and "jet" when present) and
OUT/categories.csv (geomeans of those ratios per category).
"""
import argparse
import json
import math
import os
import re
import statistics
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CLI = os.path.join(HERE, "bin", "gpart_partition_standard")
VERIFY = os.path.join(HERE, "bin", "gpart_verify_csr")

RUN_COLUMNS = ["category", "graph", "k", "repeat", "method", "status", "seconds", "reported_cut", "cut",
               "cut_fraction", "max_load_ratio", "capacity", "capacity_ok", "gpu_uuid", "binary_sha256",
               "coarsen_mode", "coarsen_seconds", "initial_partition_seconds", "uncoarsen_seconds",
               "cleanup_and_boundary_seconds"]
STAGES = ["coarsen_seconds", "initial_partition_seconds", "uncoarsen_seconds"]


def _package():
    sys.path.insert(0, ROOT)
    from tests.conftest import load_package
    return load_package()


def graph_files(spec, out):
    """Writes (or names) the int64 CSR files of one graph spec; returns
    (name, category, indptr path, indices path)."""
    kind, _, rest = spec.partition(":")
    if kind == "file":
        ip, ix = rest.split(":")
        return os.path.splitext(os.path.basename(ip))[0], "File", ip, ix
    gp = _package()
    if kind == "grid":
        r, c = (int(x) for x in rest.split("x"))
        off, nbr = gp.generate_grid(r, c)
        name, category = "grid%dx%d" % (r, c), "Artificial Mesh"
    elif kind == "rmat":
        parts = [int(x) for x in rest.split(":")]
        scale, ef, seed = parts[0], parts[1], parts[2] if len(parts) > 2 else 1
        off, nbr = gp.generate_rmat(scale, ef, 0.57, 0.19, 0.19, seed=seed)
        name, category = "rmat%d_ef%d_s%d" % (scale, ef, seed), "Artificial Complex"
    elif kind == "rgg":
        parts = [int(x) for x in rest.split(":")]
        n, seed = parts[0], parts[1] if len(parts) > 1 else 1
        off, nbr = gp.generate_rgg2d(n, (8.0 / (math.pi * n)) ** 0.5, seed=seed)
        name, category = "rgg%d_s%d" % (n, seed), "Random Geometric"
    else:
        raise ValueError("unknown graph spec " + spec)
    d = os.path.join(out, "graphs")
    os.makedirs(d, exist_ok=True)
    ip, ix = os.path.join(d, name + ".indptr"), os.path.join(d, name + ".indices")
    if not (os.path.exists(ip) and os.path.exists(ix)):
        gp.save_csr(ip, ix, off, nbr)
    return name, category, ip, ix


def parse_log(text):
    """(seconds, reported cut, stage seconds) from a CLI log, as run.sh parses it."""
    timing = re.findall(r"^partition_timing (.*)$", text, re.M)[-1]
    stage = {key: float(value) for key, value in re.findall(r"(\w+_seconds)=([0-9.eE+-]+)", timing)}
    cut = int(re.findall(r"^final_edge_cut=(\d+)$", text, re.M)[-1])
    seconds = stage["algorithm_total_seconds"]
    if not math.isfinite(seconds) or seconds <= 0 or cut < 0:
        raise ValueError("invalid algorithm time or cut")
    modes = re.findall(r"\bcoarsen_mode=(\w+)", text)
    return seconds, cut, stage, modes[-1] if modes else None


def run(args):
    out = os.path.abspath(args.out)
    for sub in ("labels", "logs"):
        os.makedirs(os.path.join(out, sub), exist_ok=True)
    methods = {"ours": CLI}
    for m in args.method or []:
        name, _, path = m.partition("=")
        methods[name] = path
    names = list(methods)
    parts = [int(x) for x in args.parts.split(",")]
    graphs = {}
    for spec in args.graph:
        name, category, ip, ix = graph_files(spec, out)
        graphs[name] = dict(category=category, indptr=ip, indices=ix)
    json.dump(dict(methods=names, parts=parts, repeats=args.repeats, graphs=graphs,
                   binaries={n: dict(path=p, sha256=_sha256(p)) for n, p in methods.items()}),
              open(os.path.join(out, "meta.json"), "w"), indent=1)
    log_path = os.path.join(out, "runs.jsonl")
    for name, g in graphs.items():
        for method in names:  # one warm-up per graph and method (run.sh:219-233)
            subprocess.run([methods[method], g["indptr"], g["indices"], str(parts[0]),
                            os.path.join(out, "labels", "warmup.part"), "1.03", "0"],
                           capture_output=True, timeout=args.timeout)
        for k in parts:
            for repeat in range(args.repeats):
                shift = repeat % len(names)
                for method in names[shift:] + names[:shift]:
                    stem = "%s-k%d-r%d-%s" % (name, k, repeat, method)
                    label = os.path.join(out, "labels", stem + ".part")
                    log = os.path.join(out, "logs", stem + ".log")
                    cmd = [methods[method], g["indptr"], g["indices"], str(k), label, "1.03", "0"]
                    started = time.time()
                    try:
                        res = subprocess.run(cmd, capture_output=True, text=True, timeout=args.timeout)
                        code, text, expired = res.returncode, res.stdout + res.stderr, False
                    except subprocess.TimeoutExpired as e:
                        code, text, expired = -9, (e.stdout or b"").decode(errors="replace"), True
                    open(log, "w").write(text)
                    row = dict(id=stem, graph=name, category=g["category"], k=k, repeat=repeat, method=method,
                               gpu_uuid=os.environ.get("GPU_UUID", "device0"), started_unix=started,
                               finished_unix=time.time(), command=cmd, partition=label, log=log, returncode=code,
                               status="timeout" if expired else "failed" if code else "ok")
                    if row["status"] == "ok":
                        try:
                            seconds, cut, stage, mode = parse_log(text)
                            row.update(seconds=seconds, reported_cut=cut, stage=stage, coarsen_mode=mode)
                        except (ValueError, KeyError, IndexError) as error:
                            row.update(status="parse_failed", error=str(error))
                    if row["status"] != "ok":
                        row["error_tail"] = text[-2000:]
                    with open(log_path, "a") as f:
                        f.write(json.dumps(row) + "\n")
                    print("%s k=%d repeat=%d %s: %s" % (name, k, repeat + 1, method, row["status"]), flush=True)


def _sha256(path):
    import hashlib
    h = hashlib.sha256()
    try:
        with open(path, "rb") as f:
            for chunk in iter(lambda: f.read(1 << 20), b""):
                h.update(chunk)
    except OSError:
        return None
    return h.hexdigest()


def verify(run_row, graph):
    """gpart_verify_csr on one partition file: exact cut and loads against
    the eps 1.03 capacity (analyze_cuts.sh's check)."""
    res = subprocess.run([VERIFY, graph["indptr"], graph["indices"], str(run_row["k"]), run_row["partition"],
                          "binary", "--symmetric"], capture_output=True, text=True)
    try:
        exact = json.loads(res.stdout)
    except ValueError:
        return dict(status="verify_failed"), {}
    ok = (res.returncode == 0 and exact.get("capacity_ok") and exact.get("cut") == run_row.get("reported_cut"))
    return dict(status="verified" if ok else "mismatch"), exact


def _write(path, rows, columns):
    import csv
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=columns, extrasaction="ignore")
        w.writeheader()
        for r in rows:
            w.writerow({c: r.get(c) for c in columns})


def _geomean(xs):
    return math.exp(statistics.mean(map(math.log, xs))) if xs else None


def analyze(args, verify_fn=verify):
    out = os.path.abspath(args.out)
    meta = json.load(open(os.path.join(out, "meta.json")))
    runs = [json.loads(line) for line in open(os.path.join(out, "runs.jsonl")) if line.strip()]
    graphs = meta["graphs"]
    rows, groups = [], {}
    for r in runs:
        check, exact = (verify_fn(r, graphs[r["graph"]]) if r["status"] == "ok" else (dict(status=r["status"]), {}))
        row = dict(category=r["category"], graph=r["graph"], k=r["k"], repeat=r["repeat"] + 1, method=r["method"],
                   status=check["status"], seconds=r.get("seconds"), reported_cut=r.get("reported_cut"),
                   cut=exact.get("cut"), cut_fraction=exact.get("cut_ratio"),
                   max_load_ratio=exact.get("max_load_ratio"), capacity=exact.get("capacity"),
                   capacity_ok=exact.get("capacity_ok"), gpu_uuid=r.get("gpu_uuid"),
                   binary_sha256=meta["binaries"].get(r["method"], {}).get("sha256"),
                   coarsen_mode=r.get("coarsen_mode"))
        stage = r.get("stage") or {}
        for field in STAGES:
            row[field] = stage.get(field)
        durations = [row[f] for f in STAGES]
        row["cleanup_and_boundary_seconds"] = (r["seconds"] - sum(durations)
                                               if r.get("seconds") is not None and all(d is not None for d in durations)
                                               else None)
        rows.append(row)
        groups.setdefault((r["graph"], r["k"], r["method"]), []).append(row)
    _write(os.path.join(out, "runs.csv"), rows, RUN_COLUMNS)

    columns = ["category", "graph", "k", "status"]
    for method in meta["methods"]:
        columns += [method + "_" + f for f in ("valid_repeats", "seconds", "cut", "cut_fraction")]
    columns += ["speedup_over_jet", "time_reduction_vs_jet_percent", "cut_ratio_over_jet", "cut_change_vs_jet_percent",
                "speedup_over_reference", "cut_ratio_over_reference", "speedup_over_legacy",
                "time_reduction_vs_legacy_percent", "cut_ratio_over_legacy", "target_status"]
    comparison = []
    for name, graph in graphs.items():
        for k in meta["parts"]:
            row = dict(category=graph["category"], graph=name, k=k, status="complete")
            for method in meta["methods"]:
                measured = groups.get((name, k, method), [])
                valid = [r for r in measured if r["status"] == "verified"]
                row[method + "_valid_repeats"] = len(valid)
                if len(measured) != meta["repeats"] or len(valid) != meta["repeats"]:
                    row["status"] = "incomplete_or_invalid"
                    continue
                for field in ("seconds", "cut", "cut_fraction"):
                    row[method + "_" + field] = statistics.median(r[field] for r in valid)
            if row["status"] == "complete":
                for other, tag in (("jet", "jet"), ("reference", "reference"), ("legacy", "legacy")):
                    if other not in meta["methods"] or "ours" not in meta["methods"]:
                        continue
                    speed = row[other + "_seconds"] / row["ours_seconds"]
                    cut, base = row["ours_cut"], row[other + "_cut"]
                    ratio = cut / base if base else (1.0 if cut == 0 else None)
                    row["speedup_over_" + tag] = speed
                    row["cut_ratio_over_" + tag] = ratio
                    if tag in ("jet", "legacy"):
                        row["time_reduction_vs_%s_percent" % tag] = 100 * (1 - 1 / speed)
                    if tag == "jet":
                        row["cut_change_vs_jet_percent"] = 100 * (ratio - 1) if ratio is not None else None
                        row["target_status"] = ("below_1.15" if speed < 1.15 else "below_1.2" if speed < 1.2
                                                else "at_least_1.2")
            comparison.append(row)
    _write(os.path.join(out, "comparison.csv"), comparison, columns)

    categories = []
    for category in dict.fromkeys(g["category"] for g in graphs.values()):
        planned = [r for r in comparison if r["category"] == category]
        complete = [r for r in planned if r["status"] == "complete"]
        row = dict(category=category, planned_configs=len(planned), complete_configs=len(complete))
        for tag in ("jet", "legacy", "reference"):
            speeds = [r["speedup_over_" + tag] for r in complete if r.get("speedup_over_" + tag) is not None]
            ratios = [r["cut_ratio_over_" + tag] for r in complete if r.get("cut_ratio_over_" + tag) is not None]
            row["speedup_over_%s_geomean" % tag] = _geomean(speeds)
            row["mean_cut_ratio_over_" + tag] = statistics.mean(ratios) if ratios else None
        row["mean_ours_cut_fraction"] = (statistics.mean(r["ours_cut_fraction"] for r in complete)
                                         if complete and "ours" in meta["methods"] else None)
        row["mean_jet_cut_fraction"] = (statistics.mean(r["jet_cut_fraction"] for r in complete)
                                        if complete and "jet" in meta["methods"] else None)
        speeds = [r["speedup_over_jet"] for r in complete if r.get("speedup_over_jet") is not None]
        row["below_1_15"] = sum(s < 1.15 for s in speeds)
        row["below_1_2"] = sum(s < 1.2 for s in speeds)
        categories.append(row)
    _write(os.path.join(out, "categories.csv"), categories,
           ["category", "planned_configs", "complete_configs", "speedup_over_jet_geomean", "mean_cut_ratio_over_jet",
            "speedup_over_legacy_geomean", "mean_cut_ratio_over_legacy", "speedup_over_reference_geomean",
            "mean_cut_ratio_over_reference", "mean_ours_cut_fraction", "mean_jet_cut_fraction", "below_1_15",
            "below_1_2"])
    bad = sum(r["status"] != "verified" for r in rows)
    print("Analysis complete: %d records; %d failed/invalid records retained. CSV: %s" % (len(rows), bad, out),
          flush=True)
    return rows, comparison, categories


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run")
    r.add_argument("out")
    r.add_argument("--graph", action="append", required=True)
    r.add_argument("--parts", default="64")
    r.add_argument("--repeats", type=int, default=3)
    r.add_argument("--method", action="append")
    r.add_argument("--timeout", type=float, default=600.0)
    a = sub.add_parser("analyze")
    a.add_argument("out")
    args = ap.parse_args(argv)
    if args.cmd == "run":
        run(args)
    else:
        analyze(args)


if __name__ == "__main__":
    main()
