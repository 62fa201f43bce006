"""The reference's prefix-sum graph (ptaskUnitTest/cpfxsum.cpp:174-299, run_cpfxsum_task) through
the PTask runtime (ptask_amd/csrc/apps/ptk_pfxsum.cpp): one `pfxsum` task, pin / pout / sticky N,
a 16-block grid of 256 threads.  Inputs are the reference's own fixtures (g_cdata, of which
cpfxsum uses the first 24 values, and g_data of pfxsum.cpp) with their expected exclusive scans
(tests/golden/pfxsum_fixtures.json), plus larger seeded arrays checked against the oracle.  On the
host the task is the oracle's host form; on the GPU, scan.hsaco's kernel.
"""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from ptask_amd import paths
from test_pca_graph import HOSTLIB

APP = paths.bin_path("ptk_pfxsum")
FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "pfxsum_fixtures.json")))


def run(kernels, tmp_path, data=None, n=None, iterations=2):
    dump = str(tmp_path / "out.bin")
    cmd = [APP, "--kernels", kernels, "--iterations", str(iterations), "--dump", dump]
    if data is not None:
        f = str(tmp_path / "in.bin")
        np.asarray(data, dtype=np.int32).tofile(f)
        cmd += ["--data", f]
    else:
        cmd += ["--n", str(n)]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr + p.stdout
    js = json.loads(p.stdout.strip().splitlines()[-1])
    return js, np.fromfile(dump, dtype=np.int32)


def cases():
    c = FIX["g_cdata"]
    u = c["used_prefix"]
    return [("g_cdata[:24]", c["input"][:u], c["exclusive_scan"][:u]),
            ("g_cdata", c["input"], c["exclusive_scan"]),
            ("g_data", FIX["g_data"]["input"], FIX["g_data"]["exclusive_scan"])]


@pytest.mark.parametrize("name,data,want", cases(), ids=[c[0] for c in cases()])
def test_host_graph_fixtures(tmp_path, name, data, want):
    js, got = run(HOSTLIB, tmp_path, data=data)
    assert js["succeeded"] and js["backend"] == "host"
    assert got.tolist() == want


@pytest.mark.gpu
@pytest.mark.parametrize("name,data,want", cases(), ids=[c[0] for c in cases()])
def test_gpu_graph_fixtures(tmp_path, name, data, want):
    print(f"case pfxsum graph fixture {name}", flush=True)
    js, got = run(paths.kernel_path("scan.hsaco"), tmp_path, data=data)
    assert js["succeeded"] and js["backend"] == "hip"
    assert got.tolist() == want


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 255, 256, 257, 100_003])
def test_gpu_graph_seeded(tmp_path, n):
    print(f"case pfxsum graph n={n}", flush=True)
    js, got = run(paths.kernel_path("scan.hsaco"), tmp_path, n=n, iterations=1)
    assert js["succeeded"]
    src = subprocess.run([APP, "--kernels", HOSTLIB, "--n", str(n), "--dump", str(tmp_path / "h.bin")],
                         capture_output=True, text=True, timeout=300)
    assert src.returncode == 0
    # the same seeded input through the host graph gives the oracle's scan
    assert np.array_equal(got, np.fromfile(str(tmp_path / "h.bin"), dtype=np.int32))


def test_oracle_agrees_with_fixture_scan():
    c = FIX["g_cdata"]
    got = oracle.pfxsum(np.asarray(c["input"], dtype=np.int32))
    assert got.tolist() == c["exclusive_scan"]
