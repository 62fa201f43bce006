"""Several FDTD graphs live on one runtime at once (ptk_fdtd --graphs): the concurrent-graph
scenario of ptaskUnitTest/rttests.cpp:936-1333 (run_concurrent_graph_test) with per-graph task
priorities as graphbench.cpp:752-845 sets them.

Every graph's inputs are pushed before any output is pulled, so the graphs' dispatches contend for
the accelerator through the runtime's admission arbiter together.  Each graph starts from its own
seed; each result must equal the oracle's fields for that seed bit for bit (graphfdtd.cpp:185-340)
and each graph must have dispatched exactly its own tasks.  On the host the tasks are the oracle's
host-task forms; on the GPU they are the HIP task kernels of fdtd.hsaco / fdtd_fused.hsaco, on one
stream per task, so kernels of different graphs overlap on the device.
"""
import subprocess

import numpy as np
import pytest

import oracle
from ptask_amd import paths
from test_runtime_fdtd_graph import APP, HOSTLIB, load_dump, run_app

TASKS_PER_STEP = {"six": 6, "halves": 2, "fused": 1}


def check_graphs(r, tmp_dump, n, steps, seed0, graphs, form):
    assert r["ok"] and r["graphs"] == graphs
    res = r["results"]
    assert [x["graph"] for x in res] == list(range(graphs))
    for g in range(graphs):
        assert res[g]["seed"] == seed0 + g
        assert res[g]["dispatches"] == TASKS_PER_STEP[form] * steps
        got = load_dump(f"{tmp_dump}.{g}", n, n, n)
        F = oracle.init_fields(n, n, n, seed0 + g)
        want = oracle.fdtd_run(F, oracle.make_params(n, n, n, 0.5), steps)
        for f in range(6):
            assert np.array_equal(got[f].view(np.uint32), want[f].view(np.uint32)), (g, oracle.FIELDS[f])
    # graphs with different seeds end in different states
    assert len({tuple(x["fnv1a"]) for x in res}) == graphs


@pytest.mark.parametrize("mode", ["looped", "unrolled"])
@pytest.mark.parametrize("sched", ["priority", "fifo"])
def test_host_concurrent_graphs(tmp_path, mode, sched):
    d = str(tmp_path / "c.bin")
    r = run_app("--kernels", HOSTLIB, "--n", 7, "--steps", 4, "--mode", mode, "--graphs", 4,
                "--priorities", "1,5,3,9", "--sched", sched, "--seed", 40, dump=d)
    assert r["backend"] == "host" and r["sched"] == sched
    assert [x["priority"] for x in r["results"]] == [1, 5, 3, 9]
    check_graphs(r, d, 7, 4, 40, 4, "six")


def test_concurrent_graphs_rejects_chain():
    p = subprocess.run([APP, "--kernels", HOSTLIB, "--graphs", "2", "--chain"], capture_output=True, text=True)
    assert p.returncode == 2


@pytest.mark.gpu
@pytest.mark.parametrize("form,kernels", [("six", "fdtd.hsaco"), ("halves", "fdtd_fused.hsaco")])
def test_gpu_concurrent_graphs_with_priorities(tmp_path, form, kernels):
    print(f"case concurrent HIP graphs: 6 x {form} looped 40^3, 12 steps", flush=True)
    d = str(tmp_path / "c.bin")
    r = run_app("--kernels", paths.kernel_path(kernels), "--n", 40, "--steps", 12, "--mode", "looped",
                "--form", form, "--graphs", 6, "--priorities", "0,10,2,8,4,6", "--seed", 300, dump=d, timeout=600)
    assert r["backend"] == "hip"
    check_graphs(r, d, 40, 12, 300, 6, form)


@pytest.mark.gpu
def test_gpu_concurrent_unrolled_graphs_fifo(tmp_path):
    print("case concurrent HIP graphs: 3 x six unrolled 33^3, 5 steps, fifo", flush=True)
    d = str(tmp_path / "c.bin")
    r = run_app("--kernels", paths.kernel_path("fdtd.hsaco"), "--n", 33, "--steps", 5, "--mode", "unrolled",
                "--graphs", 3, "--sched", "fifo", "--seed", 500, dump=d, timeout=600)
    assert r["backend"] == "hip"
    check_graphs(r, d, 33, 5, 500, 3, "six")
