"""The drop-in C++ surface: the interface headers (include/*.hpp, through the
compiled tests/cpp/api_test.cpp) and the gpart_partition_standard CLI
(app/partition_main.cu:21-183 of the reference).

CPU: the headers' host API, and the CLI's usage (2) and error (1) exits
before any device use.  GPU: the headers' device API against the C ABI, and a
CLI run of the 256x256 grid at k=8 with --verify whose stdout keys, exit code
and partition file are checked, the file also by gpart_verify_csr.
"""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "acc-gp_amd", "bin")
API_TEST = os.path.join(BIN, "gpart_api_test")
CLI = os.path.join(BIN, "gpart_partition_standard")
VERIFY = os.path.join(BIN, "gpart_verify_csr")


def _say(msg):
    print(msg, flush=True)


def _run(args, timeout=120, env=None):
    full = dict(os.environ)
    if env:
        full.update(env)
    return subprocess.run(args, capture_output=True, text=True, timeout=timeout, env=full)


def test_api_headers_host_side(tmp_path):
    res = _run([API_TEST, "cpu", str(tmp_path)])
    assert res.returncode == 0, res.stdout + res.stderr
    assert "PASSED 0" in res.stdout and "FAIL" not in res.stdout


def _grid_files(gp, tmp_path, side=16):
    off, nbr = gp.generate_grid(side, side)
    ip, ix = str(tmp_path / "g.indptr"), str(tmp_path / "g.indices")
    gp.save_csr(ip, ix, off, nbr)
    return off, nbr, ip, ix


def test_cli_usage_and_error_exits(gp, tmp_path):
    off, nbr, ip, ix = _grid_files(gp, tmp_path)
    out = str(tmp_path / "part.bin")
    # usage errors: exit 2 (app/partition_main.cu:32-36)
    for args in ([], [ip, ix, "4"], [ip, ix, "4", out, "1.03", "0", "extra"]):
        res = _run([CLI] + args)
        assert res.returncode == 2 and "Usage:" in res.stderr, args
    # exceptions: exit 1 with the message on stderr (app/partition_main.cu:179-182)
    res = _run([CLI, ip, ix, "4", out], env={"GPART_JOINT_ROUNDS": "3"})
    assert res.returncode == 1 and "GPART_JOINT_ROUNDS was retired" in res.stderr
    res = _run([CLI, ip, ix, "1", out])
    assert res.returncode == 1 and "invalid partition capacity options" in res.stderr
    res = _run([CLI, ip, ix, "4", out], env={"GPART_COARSEN_MODE": "sclp"})
    assert res.returncode == 1 and "Only regret coarsening" in res.stderr
    res = _run([CLI, ip, ix, "4", out], env={"GPART_COARSE_BETA": "1"})
    assert res.returncode == 1 and "GPART_COARSE_BETA" in res.stderr
    bad = str(tmp_path / "bad.indices")
    np.concatenate([nbr[:-1], [len(off) + 5]]).astype("<i8").tofile(bad)
    res = _run([CLI, ip, bad, "4", out])
    assert res.returncode == 1 and "invalid neighbor" in res.stderr
    assert not os.path.exists(out)


@pytest.mark.gpu
def test_api_headers_device_side(tmp_path):
    _say("case api_test gpu: C++ headers against the C ABI")
    res = _run([API_TEST, "gpu", str(tmp_path)], timeout=300)
    _say(res.stdout[-3000:])
    assert res.returncode == 0, res.stdout[-3000:] + res.stderr[-2000:]
    assert "PASSED 0" in res.stdout


@pytest.mark.gpu
def test_cli_grid256_k8_verify(gp, ctx, tmp_path):
    _say("case cli grid256 k=8 --verify")
    off, nbr, ip, ix = _grid_files(gp, tmp_path, side=256)
    out = str(tmp_path / "part.bin")
    res = _run([CLI, ip, ix, "8", out, "1.03", "0", "--verify"], timeout=300)
    assert res.returncode == 0, res.stdout[-3000:] + res.stderr[-2000:]
    lines = res.stdout.splitlines()
    keys = {}
    for line in lines:
        for tok in line.split():
            if "=" in tok:
                k, v = tok.split("=", 1)
                keys.setdefault(k, v)
    assert keys["verify"] == "1" and keys["coarsen_mode"] == "regret"
    assert any(line.startswith("ml_layer_verify level=0 ") for line in lines)
    assert any(line.startswith("partition_timing ") for line in lines)
    assert float(keys["algorithm_total_seconds"]) > 0
    cut = int(keys["final_edge_cut"])
    assert int(keys["cut_recount"]) == cut and keys["verify_status"] == "ok"
    labels = np.fromfile(out, "<i4")
    assert len(labels) == len(off) - 1
    # same labels as the library's whole pipeline in this process
    g = gp.Graph(ctx, off, nbr)
    try:
        lab, st = g.partition(8, 1.03, 0)
    finally:
        g.close()
    assert np.array_equal(labels, lab) and st["final_cut"] == cut
    chk = _run([VERIFY, ip, ix, "8", out, "binary", "--symmetric"])
    got = json.loads(chk.stdout)
    assert chk.returncode == 0 and got["cut"] == cut and got["capacity"] == 8437
    assert max(got["loads"]) <= 8437
