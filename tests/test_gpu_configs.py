"""Every measured workload at its own size on the GPU, checked against the oracle.

- FDTD 384^3: the looped three-step persistent graph (one block per compute unit, the production
  schedule of FdtdThreeStepPersistent) for 8 steps (two three-step launches and the two-step
  remainder task), and the reference's six-task graph for 2 steps; both bitwise against the
  binary32 oracle (graphfdtd.cpp:185-340).
- EM-GMM N = 2,000,000, D = 32, K = 64, three EM steps through the GMM graph: every log-likelihood
  within 1e-4 relative of the binary64 oracle (gaumixmod.h:28-73).
- PCA 65536 x 2048, K = 16, J = 2 through the PCA graph: singular values within 1e-4 relative of
  the binary64 oracle (pca_cpu.cpp:267-424), loads and scores 1e-3, residual 1e-4 of scale.
- Batched: the full mix of 2048 FDTD instances (edge 16..96, 50 steps) and 512 groupby instances
  (1M..8M rows, Zipf 1.1 over 65536 keys) in one executor; a seeded sample of 16 FDTD instances
  (the smallest and largest edges among them) and 4 groupby instances (one of 8M rows) is read
  back and compared with the oracle.
- Served: a Poisson stream through ptk_serve; every request is re-run alone and its result
  checksums must equal the in-mix ones.

The oracle's EM and PCA run in their multi-threaded forms, bitwise equal to the serial ones.
"""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from ptask_amd import batch, paths
from test_gmm_graph import run_gmm
from test_gpu_gmm import check_against_oracle
from test_pca_graph import compare, run_pca
from test_runtime_fdtd_graph import load_dump, run_app

pytestmark = pytest.mark.gpu

GB = 1 << 30


def fdtd_oracle(n, steps, seed):
    F = oracle.init_fields(n, n, n, seed)
    return oracle.fdtd_run(F, oracle.make_params(n, n, n, 0.5), steps)


def assert_fields_bitwise(got, want):
    for f in range(6):
        g, w = got[f].view(np.uint32), want[f].view(np.uint32)
        bad = np.flatnonzero(g != w)
        assert bad.size == 0, f"{oracle.FIELDS[f]}: {bad.size} cells differ, first at {bad[:4]}"


def test_fdtd_384_persistent_three_step_graph(tmp_path):
    print("case config fdtd 384^3 overlap3p 8 steps", flush=True)
    d = str(tmp_path / "g.bin")
    g = run_app("--kernels", paths.kernel_path("fdtd_fused.hsaco"), "--form", "overlap3p", "--n", 384,
                "--steps", 8, "--seed", 5, "--mode", "looped", dump=d, timeout=600)
    assert g["ok"] and g["backend"] == "hip"
    assert g["dispatches_last_run"] == 3  # 8 = 2 x 3 + the two-step remainder
    assert_fields_bitwise(load_dump(d, 384, 384, 384), fdtd_oracle(384, 8, 5))


def test_fdtd_384_six_task_graph(tmp_path):
    print("case config fdtd 384^3 six tasks 2 steps", flush=True)
    d = str(tmp_path / "g.bin")
    g = run_app("--kernels", paths.kernel_path("fdtd.hsaco"), "--form", "six", "--n", 384,
                "--steps", 2, "--seed", 6, dump=d, timeout=600)
    assert g["ok"] and g["backend"] == "hip" and g["dispatches_last_run"] == 12
    assert_fields_bitwise(load_dump(d, 384, 384, 384), fdtd_oracle(384, 2, 6))


@pytest.mark.parametrize("accumulate", ["valu", "mfma"])
def test_gmm_2m_d32_k64_graph(tmp_path, accumulate):
    print(f"case config gmm N=2M D=32 K=64 3 steps {accumulate}", flush=True)
    js, X, mu0, m = run_gmm(paths.kernel_path("gmm.hsaco"), tmp_path, 2_000_000, 32, 64, 3,
                            "--mode", "looped", "--accumulate", accumulate, timeout=600)
    assert js["ok"] and js["backend"] == "hip" and js["nonpositive_var"] == 0
    check_against_oracle(X, mu0, m, 3, f"graph N=2M D=32 K=64 {accumulate}", threads=oracle.default_threads())


def test_pca_65536x2048_k16_graph(tmp_path):
    print("case config pca 65536x2048 K=16 J=2", flush=True)
    js, X, T, P, R, L = run_pca(paths.kernel_path("pca.hsaco"), tmp_path, 65536, 2048, 16, 2, timeout=600)
    assert js["ok"] and js["backend"] == "hip"
    compare(X, T, P, R, L, 16, 2, True, {"L": 1e-4, "P": 1e-3, "T": 1e-3, "R": 1e-4},
            threads=oracle.default_threads())


def groupby_check(ex, i, nk):
    keys, vals = ex.read(i, 0, np.int32), ex.read(i, 1, np.int32)
    g, head, offsets, ok, ov = oracle.groupby(keys, vals, nk)
    groups = ex.read(i, 2, np.int64)
    G = int(groups[0])
    assert np.array_equal(groups[:3], head)
    assert np.array_equal(groups[3:3 + 3 * G].reshape(G, 3), g)
    assert np.array_equal(ex.read(i, 3, np.int64), offsets)
    out_k, out_v = ex.read(i, 4, np.int32), ex.read(i, 5, np.int32)
    assert np.array_equal(out_k, ok)
    assert np.array_equal(out_v, ov)  # rows of a group in input order
    return keys.size


def test_batched_full_mix_sampled_against_oracle(hip):
    print("case config batched 2048 fdtd + 512 groupby", flush=True)
    rng = np.random.default_rng(2048)
    edges = rng.integers(16, 97, 2048)
    edges[[0, 1]] = (16, 96)
    rows = rng.integers(1_000_000, 8_000_001, 512)
    rows[0] = 8_000_000
    fs = np.concatenate([[0, 1], rng.choice(np.arange(2, 2048), 14, replace=False)])
    gs = np.concatenate([[0], rng.choice(np.arange(1, 512), 3, replace=False)])
    arena = sum(batch.footprint(batch.FDTD, int(e), int(e), int(e), 50) for e in edges)
    arena += sum(batch.footprint(batch.GROUPBY, int(r), 65535) for r in rows)
    with batch.Executor(int(arena * 1.05) + GB, batch.BATCHED) as ex:
        fid = [ex.submit_fdtd(int(e), int(e), int(e), 50, 10_000 + j, keep=j in fs) for j, e in enumerate(edges)]
        gid = [ex.submit_groupby(int(r), 65535, 1.1, 20_000 + j, keep=j in gs) for j, r in enumerate(rows)]
        ex.drain()
        st = ex.stats()
        assert st["graph_launches"] >= 1
        for j in fs:
            e = int(edges[j])
            got = [ex.read(fid[j], f, np.float32) for f in range(6)]
            assert_fields_bitwise(got, fdtd_oracle(e, 50, 10_000 + int(j)))
        checked = [groupby_check(ex, gid[j], 65535) for j in gs]
        assert 8_000_000 in checked
    print(f"batched mix: {len(fs)} FDTD (edges {sorted(int(edges[j]) for j in fs)}) and "
          f"{len(gs)} groupby ({checked} rows) equal to the oracle", flush=True)


def test_served_stream_every_request_identical_alone(tmp_path):
    print("case config served stream, identity of every request", flush=True)
    cmd = [paths.bin_path("ptk_serve"), "--kernels", paths.KERNEL_DIR, "--qps", "40", "--requests", "48",
           "--identity", "-1", "--seed", "0x77", "--arena-gb", "48"]
    print("run:", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stderr + p.stdout
    js = json.loads(p.stdout.strip().splitlines()[-1])
    assert js["ok"] and js["requests"] == 48
    assert js["identity_checked"] == 48 and js["identity_equal"] == 48
    assert js["fdtd"] and js["gmm"] and js["groupby"]
    print(f"served: p50 {js['p50_ms']} ms, p99 {js['p99_ms']} ms, {js['graphs_per_s']} graphs/s", flush=True)
