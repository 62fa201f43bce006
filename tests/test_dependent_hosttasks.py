"""Host tasks with a dependent HIP accelerator (include/ptask/hosttask.h; Task::
BindDependentAcceleratorClass, Port::BindDependentAccelerator; reference hosttask.h:45-144,
hosttask.cpp:357-563, task.h:768), through ptask_amd/csrc/apps/ptk_depgraph.cpp.

- sgemm: the graph of ptaskUnitTest/ptaskcublas.cpp run_cublas_task -- the EX-form host task
  SGemm (libptk_hosttasks.so) gets device pointers for A, B, C and the task's dependent stream, and
  runs rocBLAS on it; C must equal a binary64 host product to binary32 accuracy (1e-5 of scale).
- scale: a legacy-form dependent task scaling a device vector in place, iterated by a meta port,
  feeding a plain host task (the block moves device -> host between them); exact in binary32.
"""
import ctypes
import json
import os
import subprocess

import pytest

from ptask_amd import paths

APP = paths.bin_path("ptk_depgraph")
DEPLIB = paths.lib_path("libptk_hosttasks.so")
SRC = os.path.join(os.path.dirname(__file__), "hosttasks", "record_tasks.c")


@pytest.fixture(scope="module")
def hostlib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("hosttasks") / "record_tasks.so")
    subprocess.run(["gcc", "-shared", "-fPIC", "-O2", "-o", out, SRC, "-lpthread"], check=True)
    return out


def run(*args, ok=True):
    cmd = [APP, *map(str, args)]
    print("run:", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    if ok:
        assert p.returncode == 0, p.stderr + p.stdout
        return json.loads(p.stdout.strip().splitlines()[-1])
    return p


def test_library_exports_the_dependent_entry_points():
    L = ctypes.CDLL(DEPLIB)
    assert hasattr(L, "SGemm") and hasattr(L, "ScaleDeviceVector")


def test_usage_errors():
    p = run("--scenario", "scale", "--kernels", DEPLIB, ok=False)
    assert p.returncode == 2 and "usage" in p.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("m,k,n,repeat", [(64, 48, 32, 3), (1, 1, 1, 1), (257, 129, 63, 2), (512, 512, 512, 2)])
def test_sgemm_dependent_ex_task(m, k, n, repeat):
    r = run("--kernels", DEPLIB, "--scenario", "sgemm", "--m", m, "--k", k, "--n", n, "--repeat", repeat)
    assert r["ok"] and r["max_rel_err"] <= 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("elems,iterations", [(4096, 3), (1_000_003, 5)])
def test_legacy_dependent_task_feeds_a_host_task(hostlib, elems, iterations):
    r = run("--kernels", DEPLIB, "--host-kernels", hostlib, "--scenario", "scale", "--elems", elems,
            "--iterations", iterations, "--scalar", 1.5)
    assert r["ok"] and r["max_rel_err"] == 0
