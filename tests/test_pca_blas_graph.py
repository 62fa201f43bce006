"""The reference's unrolled binary64 PCA graph (ptaskUnitTest/graphcupca.cpp:245-1640) through the
PTask runtime (ptask_amd/csrc/apps/ptk_cupca.cpp): one task per BLAS call of pca_cpu's
Gram-Schmidt loop, bound by name to ptaskDcopy / ptaskDaxpy / ptaskDgemv / ptaskDnrm2 /
ptaskDscal / ptaskDger with the by-value parameter structs of graphcupca.cpp:161-220.

Every task walks its dot products and norms in pca_cpu's index order, so T, P, R and L must equal
the binary64 oracle (oracle.pca_gs, pinned against numpy's SVD in test_oracle_pca.py) bit for bit:
on the host through the oracle's host-task forms, and on the GPU through the HIP kernels of
pca.hsaco at the reference's test shape (512 x 512, K = 3, J = 10: runcupca.bat:5-7), with and
without the row centring the reference's --prenorm flag skips.
"""
import json
import subprocess

import numpy as np
import pytest

import oracle
from ptask_amd import paths
from test_pca_graph import HOSTLIB

APP = paths.bin_path("ptk_cupca")


def tasks_expected(N, K, J, prenorm):
    """Tasks of graphcupca's BuildGraph: centring (1 + (N - 1) + N), per component a copy, J
    iterations of 6 tasks (8 for k > 0) and the rank-one update, then K final scalings."""
    centring = 0 if prenorm else 2 * N
    per_k = sum(1 + J * (6 + (2 if k > 0 else 0)) + 1 for k in range(K))
    return centring + per_k + K


def run_cupca(kernels, tmp_path, M, N, K, J, *extra, timeout=300):
    dump = str(tmp_path / f"cupca_{M}_{N}_{K}_{J}.bin")
    cmd = list(map(str, [APP, "--kernels", kernels, "--m", M, "--n", N, "--k", K, "--j", J,
                         "--dump", dump, *extra]))
    print("run:", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    assert p.returncode == 0, p.stderr + p.stdout
    js = json.loads(p.stdout.strip().splitlines()[-1])
    d = np.fromfile(dump, dtype=np.float64)
    sizes = [M * N, M * K, N * K, M * N, K]
    assert d.size == sum(sizes)
    parts, o = [], 0
    for s in sizes:
        parts.append(d[o:o + s])
        o += s
    X = parts[0].reshape(N, M).T
    T = parts[1].reshape(K, M).T
    P = parts[2].reshape(K, N).T
    R = parts[3].reshape(N, M).T
    return js, X, T, P, R, parts[4]


def assert_bitwise_oracle(X, T, P, R, L, K, J, center):
    eT, eP, eR, eL = oracle.pca_gs(X, K, J, center, threads=oracle.default_threads())
    for name, got, want in (("T", T, eT), ("P", P, eP), ("R", R, eR), ("L", L, eL)):
        g = np.ascontiguousarray(got).view(np.uint64)
        w = np.ascontiguousarray(want).view(np.uint64)
        bad = np.flatnonzero(g != w)
        assert bad.size == 0, f"{name}: {bad.size} values differ, max |diff| {np.abs(got - want).max():.3e}"


def test_tasks_expected_formula():
    # graphcupca.cpp with K = 3, J = 10, N = 512: 1024 centring tasks, 3 copies, 10 x (6 + 8 + 8)
    # iteration tasks, 3 rank-one updates and 3 final scalings
    assert tasks_expected(512, 3, 10, False) == 1024 + 3 + 220 + 3 + 3


@pytest.mark.parametrize("prenorm", [False, True])
@pytest.mark.parametrize("M,N,K,J", [(40, 24, 3, 4), (17, 33, 2, 3), (1, 5, 1, 2)])
def test_host_graph_bitwise(tmp_path, M, N, K, J, prenorm):
    extra = ["--prenorm"] if prenorm else []
    js, X, T, P, R, L = run_cupca(HOSTLIB, tmp_path, M, N, K, J, *extra)
    assert js["ok"] and js["backend"] == "host"
    assert js["tasks"] == tasks_expected(N, K, J, prenorm)
    assert_bitwise_oracle(X, T, P, R, L, K, J, not prenorm)


def test_host_graph_repeated_runs(tmp_path):
    # the parameter blocks stay on their sticky ports: a second push of the inputs re-runs the
    # whole graph to the same result
    js, X, T, P, R, L = run_cupca(HOSTLIB, tmp_path, 24, 16, 2, 3, "--runs", 3, "--dist", "uniform")
    assert js["ok"]
    assert_bitwise_oracle(X, T, P, R, L, 2, 3, True)


def test_bad_arguments_rejected():
    p = subprocess.run([APP, "--kernels", HOSTLIB, "--k", "9", "--n", "4"], capture_output=True, text=True)
    assert p.returncode == 2 and "usage" in p.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("prenorm", [False, True])
def test_gpu_graph_reference_shape_bitwise(tmp_path, prenorm):
    print(f"case cupca 512x512 K=3 J=10 prenorm={prenorm}", flush=True)
    extra = ["--prenorm"] if prenorm else []
    js, X, T, P, R, L = run_cupca(paths.kernel_path("pca.hsaco"), tmp_path, 512, 512, 3, 10, *extra, timeout=600)
    assert js["ok"] and js["backend"] == "hip"
    assert js["tasks"] == tasks_expected(512, 3, 10, prenorm)
    assert_bitwise_oracle(X, T, P, R, L, 3, 10, not prenorm)


@pytest.mark.gpu
def test_gpu_graph_rectangular_bitwise(tmp_path):
    print("case cupca 300x129 K=4 J=3 uniform", flush=True)
    js, X, T, P, R, L = run_cupca(paths.kernel_path("pca.hsaco"), tmp_path, 300, 129, 4, 3, "--dist", "uniform",
                                  "--runs", 2, timeout=600)
    assert js["ok"] and js["backend"] == "hip"
    assert_bitwise_oracle(X, T, P, R, L, 4, 3, True)
