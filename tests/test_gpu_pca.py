"""Parity of the Gram-Schmidt PCA HIP task kernels (pca.hsaco) with the binary64 oracle, through
the PCA graph application.

Bar: singular values L_k within 1e-4 relative of the oracle on the same binary32 input; loads P
and scores T within 1e-3 (scores relative to their largest magnitude), the residual R within
1e-4 of the input's scale.  Errors against the binary64 run are printed.  All reductions are in a
fixed order, so a rerun is bitwise identical.
"""
import numpy as np
import pytest

from ptask_amd import paths
from test_pca_graph import compare, run_pca

pytestmark = pytest.mark.gpu

KERNELS = paths.kernel_path("pca.hsaco")
GPU_TOL = {"L": 1e-4, "P": 1e-3, "T": 1e-3, "R": 1e-4}


@pytest.mark.parametrize("M,N,K,J,dist,center", [
    (512, 512, 3, 10, "uniform", True),      # the reference's runcupca shape (K=3, J=10)
    (1000, 300, 5, 6, "normal", True),       # ragged rows and columns
    (4096, 2048, 4, 5, "normal", True),      # 2048 columns: two float4 chunks per thread
    (2000, 4096, 2, 3, "normal", False),     # 4096 columns: four chunks, 4-row slabs
    (333, 1000, 3, 4, "uniform", True),      # one chunk per thread
    (130, 3000, 3, 3, "normal", True),       # three chunks per thread
    (64, 8, 7, 4, "uniform", True),          # K = N - 1
])
def test_graph_matches_oracle(tmp_path, M, N, K, J, dist, center):
    print(f"case pca M={M} N={N} K={K} J={J} {dist} center={center}", flush=True)
    extra = ["--dist", dist] + ([] if center else ["--no-center"])
    js, X, T, P, R, L = run_pca(KERNELS, tmp_path, M, N, K, J, *extra)
    assert js["ok"] and js["backend"] == "hip"
    compare(X, T, P, R, L, K, J, center, GPU_TOL)


def test_block_counts_agree_and_reruns_are_bitwise(tmp_path):
    print("case pca blocks/determinism", flush=True)
    a = run_pca(KERNELS, tmp_path, 3000, 700, 3, 4)
    b = run_pca(KERNELS, tmp_path, 3000, 700, 3, 4, "--repeat", "2", "--warmup", "1")
    c = run_pca(KERNELS, tmp_path, 3000, 700, 3, 4, "--blocks", "37", "--gr", "5")
    assert a[0]["T_fnv1a"] == b[0]["T_fnv1a"] and a[0]["S_fnv1a"] == b[0]["S_fnv1a"]
    assert b[0]["h2d_count_timed"] == 0
    assert np.allclose(a[5], c[5], rtol=1e-5)
    compare(c[1], c[2], c[3], c[4], c[5], 3, 4, True, GPU_TOL)


def test_captured_runs_equal_per_task_dispatch(tmp_path):
    # Graph::EnableLaunchCapture: the first run is recorded, the later ones replay its launches
    # as one hipGraph; the last run's T, S and R must equal per-task dispatch bit for bit
    print("case pca capture", flush=True)
    args = ("--repeat", "3", "--warmup", "1")
    a = run_pca(KERNELS, tmp_path, 3000, 700, 3, 4, *args)
    b = run_pca(KERNELS, tmp_path, 3000, 700, 3, 4, *args, "--capture")
    assert not a[0]["capture"]["armed"]
    assert b[0]["capture"]["armed"] and b[0]["capture"]["replays"] == 3
    assert a[0]["T_fnv1a"] == b[0]["T_fnv1a"] and a[0]["S_fnv1a"] == b[0]["S_fnv1a"]
    assert np.array_equal(a[4], b[4])
