"""The EM-GMM graph (ptask_amd/csrc/apps/ptk_gmm.cpp) on the HostAccelerator path.

The host-task forms of GmmAccumulate/GmmFinalize (oracle/gmm.c) run through the PTask runtime in
both graph forms (looped with general iterators, unrolled); the final model must equal the
binary64 oracle EM on the same points to rounding (different summation order only).
"""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from ptask_amd import gmm, paths

APP = paths.bin_path("ptk_gmm")
HOSTLIB = os.path.join(oracle.ORACLE_DIR, "_build", "libptk_oracle.so")


def run_gmm(kernels, tmp_path, N, D, K, steps, *extra, timeout=300):
    dump = str(tmp_path / f"gmm_{N}_{D}_{K}_{steps}.bin")
    cmd = [APP, "--kernels", kernels, "--n", N, "--d", D, "--k", K, "--steps", steps,
           "--dump", dump, *extra]
    cmd = list(map(str, cmd))
    print("run:", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    assert p.returncode == 0, p.stderr + p.stdout
    js = json.loads(p.stdout.strip().splitlines()[-1])
    X, mu0, model = gmm.read_dump(dump, N, D, K, steps)
    return js, X, mu0, model


def oracle_em(X, mu0, steps):
    return oracle.gmm_em(X.astype(np.float64), mu0, steps - 1)


@pytest.mark.parametrize("mode", ["looped", "unrolled"])
@pytest.mark.parametrize("N,D,K,steps", [(3000, 4, 3, 6), (1000, 1, 2, 8), (517, 3, 5, 4)])
def test_host_graph_equals_oracle(tmp_path, mode, N, D, K, steps):
    js, X, mu0, m = run_gmm(HOSTLIB, tmp_path, N, D, K, steps, "--mode", mode, "--init", "strided")
    assert js["ok"] and js["backend"] == "host" and m.steps == steps
    mu, var, frac, ll = oracle_em(X, mu0, steps)
    assert np.allclose(m.mu, mu, rtol=1e-10, atol=1e-10)
    assert np.allclose(m.var, var, rtol=1e-8, atol=1e-12)
    assert np.allclose(m.frac, frac, rtol=1e-10, atol=1e-14)
    assert np.allclose(m.loglike, ll, rtol=1e-10)


def test_component_init_converges_without_collapse(tmp_path):
    """Generated points with one initial mean per generating component (the default): EM keeps
    every component populated and the log-likelihood non-decreasing."""
    js, X, mu0, m = run_gmm(HOSTLIB, tmp_path, 20000, 8, 12, 6)
    assert js["init"] == "components" and js["nonpositive_var"] == 0
    assert np.all(m.frac > 0.02) and abs(m.frac.sum() - 1) < 1e-12
    assert np.all(np.diff(m.loglike[1:]) >= -1e-9 * np.abs(m.loglike[1:-1]))
    mu, var, frac, ll = oracle_em(X, mu0, 6)
    assert np.allclose(m.loglike, ll, rtol=1e-10)


def test_host_graph_forms_and_reruns_agree(tmp_path):
    a = run_gmm(HOSTLIB, tmp_path, 2000, 2, 4, 5, "--mode", "looped", "--repeat", "3",
                "--warmup", "1")[0]
    b = run_gmm(HOSTLIB, tmp_path, 2000, 2, 4, 5, "--mode", "unrolled")[0]
    assert a["model_fnv1a"] == b["model_fnv1a"]
    assert a["h2d_count_timed"] == 0 and a["steps_done"] == 5


def test_host_graph_on_reference_fixture(tmp_path):
    """dist_3 (the reference's R draws, 1-D) through the graph; the initial means are the points
    at rows c*N/K, as the app chooses them."""
    x = np.loadtxt(os.path.join(os.path.dirname(__file__), "golden", "dist_3.txt"))
    x32 = x.astype(np.float32)
    path = tmp_path / "dist3.f32"
    x32.tofile(path)
    js, X, mu0, m = run_gmm(HOSTLIB, tmp_path, x.size, 1, 3, 21, "--data", str(path))
    mu, var, frac, ll = oracle_em(X, mu0, 21)
    assert np.allclose(m.mu, mu, rtol=1e-10) and np.allclose(m.loglike, ll, rtol=1e-10)
    assert np.all(np.diff(m.loglike[1:]) >= -1e-9 * np.abs(m.loglike[1:-1]))


@pytest.mark.parametrize("args", [["--d", "33"], ["--k", "65"], ["--k", "0"], ["--steps", "0"],
                                  ["--mode", "bogus"], ["--n", "3", "--k", "4"],
                                  ["--init", "bogus"], ["--init", "components", "--data", "x"]])
def test_app_rejects_bad_shapes(args):
    cmd = [APP, "--kernels", HOSTLIB, "--n", "100", *args]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=60)
    assert p.returncode == 2 and "usage" in p.stderr


def test_python_model_layout_drives_host_task_forms():
    """ptask_amd/gmm.py's model layout and table (used to drive the device kernels) run through
    the host-task forms reproduce the oracle EM, and its table equals the C header's."""
    import ctypes as C
    L = oracle.lib()
    rng = np.random.default_rng(11)
    N, D, K, steps, blocks = 700, 5, 4, 5, 3
    X = (rng.uniform(-10, 10, (K, D))[rng.integers(0, K, N)]
         + rng.standard_normal((N, D))).astype(np.float32)
    m = np.zeros(gmm.model_elems(D, K, steps))
    o = gmm.model_offsets(D, K, steps)
    mu0 = X[(np.arange(K) * N) // K].astype(np.float64)
    m[1:4] = (N, D, K)
    m[o["mu"]:o["var"]] = mu0.ravel()
    m[o["var"]:o["frac"]] = 1e-10
    m[o["frac"]:o["ll"]] = 1.0 / K
    gmm.fill_table(m, D, K, steps)
    dims = gmm.Dims(N, D, K, blocks, steps)
    part = np.zeros(blocks * gmm.partial_stride(D, K))
    argv = (C.c_void_p * 4)()
    for _ in range(steps):
        argv[:] = [X.ctypes.data, C.addressof(dims), m.ctypes.data, part.ctypes.data]
        L.GmmAccumulate(C.c_uint(4), argv)
        argv[:] = [m.ctypes.data, part.ctypes.data, C.addressof(dims), None]
        L.GmmFinalize(C.c_uint(3), argv)
    got = gmm.parse_model(m, D, K, steps)
    mu, var, frac, ll = oracle.gmm_em(X.astype(np.float64), mu0, steps - 1)
    assert got.steps == steps
    assert np.allclose(got.mu, mu, rtol=1e-10, atol=1e-10)
    assert np.allclose(got.var, var, rtol=1e-8)
    assert np.allclose(got.loglike, ll, rtol=1e-10)
    # the table the C side rewrote equals the one gmm.fill_table builds from the same model
    again = m.copy()
    gmm.fill_table(again, D, K, steps)
    t0 = m[o["table"]:].view(np.float32)
    t1 = again[o["table"]:].view(np.float32)
    assert np.array_equal(t0[:2 * K * gmm.dpad(D)], t1[:2 * K * gmm.dpad(D)])
    assert np.allclose(t0[2 * K * gmm.dpad(D):gmm.table_floats(D, K)],
                       t1[2 * K * gmm.dpad(D):gmm.table_floats(D, K)], rtol=1e-6)
