"""The batched and served forms on their full workloads: every result must
equal the same request run alone, bit for bit (deterministic mode is the
only mode).

- Batched: the 256-graph mixed-size workload (tests/workloads.py, seed 5:
  grids, R-MAT and random geometric graphs, k in {4..64}) through
  gp_partition_batch against gp_partition per graph.
- Served: a Poisson stream over the same mix (k in [2, 128], every tenth
  request a repartition of a resident graph after a 1% edge delta applied
  on the device, warm-started from its current partition) submitted to a
  gp_server while earlier requests are still on the device, against
  gp_partition / gp_refine_partition alone on the host-updated CSR; the
  resident graphs and partitions at the end equal the host's.
"""
import time

import numpy as np
import pytest

from tests import workloads

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def mix(gp):
    return workloads.batch_mix(gp, 256)


def _alone(gp, ctx, off, nbr, k):
    g = gp.Graph(ctx, off, nbr)
    try:
        return g.partition(k)
    finally:
        g.close()


def test_batched_mix_256_equals_alone(gp, ctx, mix):
    assert len(mix) == 256
    _say("batched mix: %d graphs, %d undirected edges" % (len(mix), sum(len(x[2]) // 2 for x in mix)))
    items = [gp.BatchItem(off, nbr, k) for _, off, nbr, k in mix]
    stats = gp.partition_batch(items, workers=6)
    _say("batched mix: batch done; alone runs")
    for (name, off, nbr, k), it, st in zip(mix, items, stats):
        lab, ast = _alone(gp, ctx, off, nbr, k)
        assert np.array_equal(it.labels, lab), name
        assert st["final_cut"] == ast["final_cut"], name
        assert st["max_part_weight"] <= st["capacity"], name


def test_poisson_stream_equals_alone(gp, ctx, mix):
    reqs = workloads.serving_requests(gp, mix, 64)
    srv = gp.Server(workers=6)
    try:
        # Graphs that get repartitioned are resident, with the cold partition
        # at their mix k as their current partition.
        resident, host = {}, {}
        for kind, j, _, _ in reqs:
            if kind == "warm" and j not in resident:
                _, off, nbr, k0 = mix[j]
                resident[j] = srv.add_graph(off, nbr)
                srv.wait(srv.submit(gp.ResidentItem(resident[j], len(off) - 1, k0)))
                host[j] = (off, nbr, None, srv.graph(resident[j])[3])
        items = []
        for kind, j, payload, k in reqs:
            if kind == "warm":
                items.append(gp.ResidentItem(resident[j], len(mix[j][1]) - 1, k, delta=payload + (None,), warm=True))
            else:
                items.append(gp.BatchItem(payload[0], payload[1], k))
        arrivals = workloads.poisson_arrivals(len(reqs), rate=40.0)
        _say("stream: %d requests (%d warm on %d resident graphs) over %.2f s" % (
            len(reqs), sum(r[0] == "warm" for r in reqs), len(resident), arrivals[-1]))
        tickets = []
        t0 = time.perf_counter()
        for item, at in zip(items, arrivals):
            delay = t0 + at - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets.append(srv.submit(item))
        stats = [srv.wait(t) for t in tickets]
        final = {j: srv.graph(g) for j, g in resident.items()}
    finally:
        srv.close()
    lat = sorted(st["queue_seconds"] + st["device_seconds"] for st in stats)
    _say("stream: served; p50 %.1f ms p99 %.1f ms; alone runs" % (1e3 * lat[len(lat) // 2], 1e3 * lat[-1]))
    for (kind, j, payload, k), item, st in zip(reqs, items, stats):
        if kind == "cold":
            lab, ast = _alone(gp, ctx, payload[0], payload[1], k)
        else:
            off, nbr, ew, prior = host[j]
            u, v, ins = payload
            off, nbr, ew = workloads.apply_delta(off, nbr, ew, u, v, ins)
            g = gp.Graph(ctx, off, nbr, ew)
            try:
                lab, ast = g.refine(prior, gp.refine_options(k))
            finally:
                g.close()
            host[j] = (off, nbr, ew, lab)
        assert np.array_equal(item.labels, lab), (kind, j, k)
        assert st["final_cut"] == ast["final_cut"], (kind, j, k)
    for j, (off, nbr, ew, labels) in final.items():
        assert np.array_equal(off, host[j][0]) and np.array_equal(nbr, host[j][1]), j
        assert np.array_equal(labels, host[j][3]), j
