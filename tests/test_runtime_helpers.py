"""The Runtime:: helper calls of PTaskRuntime.h:2441-2560 (SetPortMetaFunction, AddTask,
BindDerivedPort, RunGraph, Push, Pull, CheckGraphSemantics, GetPhysicalAcceleratorCount,
CheckInvariantCondition) through the simple-iteration graph of ptaskUnitTest/iteration.cpp
(run_simple_iteration_task, :67-180), rebuilt with the reference test's call sequence in
ptask_amd/csrc/apps/ptk_iteration.cpp.

The task scales a vector in place; its MF_SIMPLE_ITERATOR meta port repeats the dispatch
`iterations` times and the element count arrives through a port derived from the vector
(DF_SIZE).  The result must equal A * s^iterations computed in binary32 on the host, exactly."""
import json
import os
import subprocess

import pytest

from ptask_amd import paths

APP = paths.bin_path("ptk_iteration")
SRC = os.path.join(os.path.dirname(__file__), "hosttasks", "record_tasks.c")


@pytest.fixture(scope="module")
def hostlib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("hosttasks") / "record_tasks.so")
    subprocess.run(["gcc", "-shared", "-fPIC", "-O2", "-o", out, SRC, "-lpthread"], check=True)
    return out


def run(*args):
    cmd = [APP, *map(str, args)]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr + p.stdout
    return json.loads(p.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("n,iterations,scalar", [(1000, 4, 2.0), (1, 1, 3.0), (4097, 9, -1.5), (256, 20, 0.5)])
def test_simple_iteration_graph_on_host(hostlib, n, iterations, scalar):
    r = run("--kernels", hostlib, "--n", n, "--iterations", iterations, "--scalar", scalar)
    assert r["ok"] and r["max_rel_diff"] == 0 and r["n"] == n and r["iterations"] == iterations
    assert r["semantics"] == 0 and r["invariant"]


def test_bad_kernel_name_fails_loudly(hostlib):
    p = subprocess.run([APP, "--kernels", hostlib, "--op", "NoSuchTask"], capture_output=True, text=True,
                       timeout=60)
    assert p.returncode != 0 and "GetCompiledKernel" in p.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("n,iterations", [(1000, 4), (1 << 20, 7)])
def test_simple_iteration_graph_on_hip(n, iterations):
    print(f"case iteration graph hip n={n} iterations={iterations}", flush=True)
    r = run("--kernels", paths.kernel_path("vector.hsaco"), "--n", n, "--iterations", iterations)
    assert r["ok"] and r["max_rel_diff"] == 0 and r["physical_accelerators"] >= 1
