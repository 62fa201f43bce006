"""The host runtime under ThreadSanitizer (ptask_amd/csrc/tsan.mk: libptask.so and the applications
built with -fsanitize=thread).  Scenarios that drive its threads: runner threads over looped and
unrolled host-task graphs, several graphs at once through the dispatch arbiter with priorities,
pool sweeps after forced allocation failures while runners recycle blocks, global quiescence and
stopped graphs, a block fanned out to an in/out writer and a reader (ptk_multigraph), the
Runtime:: helper sequence (ptk_iteration), and the BLAS-task PCA graph, whose buffers fan out to
readers and in-place writers at once (ptk_cupca).  A run passes
when ThreadSanitizer reports nothing beyond the suppressions of ptask_amd/csrc/tsan.supp (the
internal mutex of libstdc++'s condition_variable_any) and the results are right.
"""
import json
import os
import subprocess

import pytest

from ptask_amd import paths
from test_pca_graph import HOSTLIB as PCA_HOSTLIB
from test_runtime_fdtd_graph import HOSTLIB

TSAN_BIN = os.path.join(paths.LIB_DIR, "tsan", "bin")
SUPP = os.path.join(os.path.dirname(paths.LIB_DIR), "csrc", "tsan.supp")
RECORD_SRC = os.path.join(os.path.dirname(__file__), "hosttasks", "record_tasks.c")

pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(TSAN_BIN, "ptk_fdtd")),
                                reason="the ThreadSanitizer build (make -f tsan.mk) is not present")


def run_tsan(app, *args):
    env = dict(os.environ, TSAN_OPTIONS=f"suppressions={SUPP} second_deadlock_stack=1 halt_on_error=0")
    p = subprocess.run([os.path.join(TSAN_BIN, app), *map(str, args)], capture_output=True, text=True,
                       timeout=300, env=env)
    reports = p.stderr.count("WARNING: ThreadSanitizer")
    assert reports == 0, p.stderr[-6000:]
    return p


@pytest.fixture(scope="module")
def recordlib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("hosttasks") / "record_tasks.so")
    subprocess.run(["gcc", "-O2", "-shared", "-fPIC", "-o", out, RECORD_SRC, "-lpthread"], check=True)
    return out


@pytest.mark.parametrize("args", [
    ["--n", 8, "--steps", 4, "--repeat", 3, "--threads", 4],
    ["--n", 8, "--steps", 4, "--mode", "unrolled", "--blocks", 3, "--threads", 4],
    ["--n", 8, "--steps", 3, "--graphs", 4, "--priorities", "1,2,3,4"],
    ["--n", 8, "--steps", 3, "--repeat", 4, "--pool", "global", "--inject-host-failures", 2, "--threads", 4],
    ["--n", 8, "--steps", 3, "--repeat", 3, "--pool", "graph", "--threads", 4],
], ids=["looped", "unrolled", "concurrent-graphs", "pool-sweeps", "graph-pool"])
def test_fdtd_host_graphs_race_free(args):
    p = run_tsan("ptk_fdtd", "--kernels", HOSTLIB, *args)
    assert p.returncode == 0
    js = json.loads(p.stdout.strip().splitlines()[-1])
    assert js["ok"]


def test_fdtd_tsan_results_equal_regular_build():
    args = ["--kernels", HOSTLIB, "--n", 8, "--steps", 4, "--repeat", 2, "--threads", 4]
    a = json.loads(run_tsan("ptk_fdtd", *args).stdout.strip().splitlines()[-1])
    b = subprocess.run([paths.bin_path("ptk_fdtd"), *map(str, args)], capture_output=True, text=True, timeout=120)
    b = json.loads(b.stdout.strip().splitlines()[-1])
    assert [c["fnv1a"] for c in a["checksums"]] == [c["fnv1a"] for c in b["checksums"]]


@pytest.mark.parametrize("args", [
    ["--scenario", "contend", "--mode", "priority", "--graphs", 3, "--tasks", 3, "--priorities", "1,10,5",
     "--host-slots", 1],
    ["--scenario", "pause"],
    ["--scenario", "concurrent", "--graphs", 6, "--iterations", 5, "--host-slots", 3],
    ["--scenario", "stall", "--watchdog-ms", 50],
    ["--scenario", "fanout", "--iterations", 6, "--elems", 4096, "--host-slots", 4],
], ids=["contend", "pause", "concurrent", "watchdog", "fanout"])
def test_multigraph_race_free(recordlib, args):
    p = run_tsan("ptk_multigraph", "--kernels", recordlib, *args)
    assert p.returncode == 0
    assert json.loads(p.stdout.strip().splitlines()[-1])["ok"]


def test_runtime_helpers_race_free(recordlib):
    p = run_tsan("ptk_iteration", "--kernels", recordlib, "--iterations", 7)
    assert p.returncode == 0
    assert json.loads(p.stdout.strip().splitlines()[-1])["ok"]


def test_blas_pca_host_graph_race_free(tmp_path):
    p = run_tsan("ptk_cupca", "--kernels", PCA_HOSTLIB, "--m", 40, "--n", 24, "--k", 2, "--j", 3, "--runs", 2,
                 "--dump", tmp_path / "cupca.bin")
    assert p.returncode == 0
    assert json.loads(p.stdout.strip().splitlines()[-1])["ok"]
