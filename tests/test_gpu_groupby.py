"""Parity of the prefix-sum and groupby HIP task kernels with the oracle.

Bar: bit-exact for every integer output -- scans, per-key int64 count and sum, offsets, the
reshuffled keys and values (a group holds its rows in input order, as the oracle's stable
counting sort does), and the coalesced layout.
"""
import json
import subprocess

import numpy as np
import pytest

import oracle
from ptask_amd import paths
from test_oracle_groupby import APP, HOSTLIB, load_groupby_dump

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def smod(hip):
    return hip.Module(paths.kernel_path("scan.hsaco"))


def dev(hip, a):
    b = hip.DeviceBuffer(max(a.nbytes, 8))
    if a.nbytes:
        b.upload(a)
    return b


@pytest.mark.parametrize("n", [0, 1, 2, 24, 128, 255, 256, 257, 2047, 2048, 65539, 1 << 20])
@pytest.mark.parametrize("block", [64, 256, 1024])
def test_pfxsum_single_task_kernel(hip, smod, n, block):
    print(f"case pfxsum n={n} block={block}", flush=True)
    rng = np.random.default_rng(n)
    a = rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)
    din, dout = dev(hip, a), hip.DeviceBuffer(max(4 * n, 8))
    args = hip.ArgPack().ptr(din.ptr).ptr(dout.ptr).i32(n)
    hip.launch(smod.function("pfxsum"), (16, 1, 1), (block, 1, 1), args)  # cpfxsum's 16 blocks
    hip.synchronize()
    got = np.empty(n, np.int32)
    if n:
        dout.download(got)
    assert np.array_equal(got, oracle.pfxsum(a))


@pytest.mark.parametrize("n", [1, 4095, 4096, 4097, 1 << 20, 3_000_017])
@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_two_task_large_scan(hip, smod, n, dtype):
    print(f"case scan2 n={n} {dtype.__name__}", flush=True)
    rng = np.random.default_rng(7)
    a = rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(dtype)
    tiles = -(-n // 4096)
    sfx = "64" if dtype == np.int64 else ""
    din, dout = dev(hip, a), hip.DeviceBuffer(a.nbytes)
    dt = hip.DeviceBuffer(tiles * a.itemsize)
    for grid in (min(tiles, 2048), 3):
        hip.launch(smod.function("PfxScanReduce" + sfx), (grid, 1, 1), (1024, 1, 1),
                   hip.ArgPack().ptr(din.ptr).i32(n).ptr(dt.ptr))
        hip.launch(smod.function("PfxScanApply" + sfx), (grid, 1, 1), (1024, 1, 1),
                   hip.ArgPack().ptr(din.ptr).i32(n).ptr(dt.ptr).ptr(dout.ptr))
        hip.synchronize()
        got = np.empty_like(a)
        dout.download(got)
        assert np.array_equal(got, oracle.pfxsum(a))


def run_app(*args, timeout=300):
    cmd = [APP, *map(str, args)]
    print("run:", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    assert p.returncode == 0, p.stderr + p.stdout
    return json.loads(p.stdout.strip().splitlines()[-1])


def check_groupby(keys, vals, nk, groups, ok, ov):
    g, head, offsets, eok, eov = oracle.groupby(keys, vals, nk)
    assert np.array_equal(groups[:3], head)
    assert np.array_equal(groups[3:3 + g.size], g.reshape(-1))
    assert np.array_equal(ok, eok)
    assert np.array_equal(ov, eov)  # rows of a group in input order


@pytest.mark.parametrize("dist,n,nk", [("zipf", 200_000, 65535), ("uniform", 100_000, 31),
                                       ("zipf", 2_000_000, 65535), ("uniform", 1024, 32),
                                       ("zipf", 1, 0), ("uniform", 5000, 4999)])
def test_hip_groupby_graph_equals_oracle(tmp_path, dist, n, nk):
    d = str(tmp_path / "g.bin")
    r = run_app("--kernels", paths.kernel_path("groupby.hsaco"), "--rows", n, "--num-keys", nk,
                "--dist", dist, "--repeat", 2, "--warmup", 1, "--dump", d)
    assert r["ok"] and r["backend"] == "hip"
    keys, vals, groups, ok, ov = load_groupby_dump(d, n, nk)
    check_groupby(keys, vals, nk, groups, ok, ov)
    assert r["h2d_count_timed"] == 0  # after the warmup run the inputs stay resident


def test_hip_groupby_8m_rows_zipf_matches_host_graph():
    """The batched-config size: 8M rows, Zipf s=1.1 over 64K keys; groups and reshuffled keys and
    values bit-identical to the host graph's."""
    common = ["--rows", 8_000_000, "--keys", 65536, "--dist", "zipf", "--zipf-s", 1.1]
    g = run_app("--kernels", paths.kernel_path("groupby.hsaco"), *common)
    h = run_app("--kernels", HOSTLIB, *common)
    assert g["groups"] == h["groups"] > 60000
    assert g["outputs"][0]["fnv1a"] == h["outputs"][0]["fnv1a"]  # groups
    assert g["outputs"][1]["fnv1a"] == h["outputs"][1]["fnv1a"]  # reshuffled keys
    assert g["outputs"][2]["fnv1a"] == h["outputs"][2]["fnv1a"]  # reshuffled values


@pytest.mark.parametrize("workload", ["pfxsum", "scan2"])
def test_hip_scan_graphs_equal_oracle(tmp_path, workload):
    for n in (24, 4096, 1_000_003):
        if workload == "pfxsum" and n > 100_000:
            continue
        d = str(tmp_path / f"{workload}{n}.bin")
        run_app("--kernels", paths.kernel_path("scan.hsaco"), "--workload", workload, "--rows", n,
                "--dump", d)
        raw = np.fromfile(d, dtype=np.int32)
        assert np.array_equal(raw[2 * n:3 * n], oracle.pfxsum(raw[:n])), (workload, n)


@pytest.mark.parametrize("dist,n,nk", [("zipf", 300_001, 65535), ("uniform", 5000, 99), ("zipf", 1, 0)])
def test_hip_coalesce_equals_reference_layout(tmp_path, dist, n, nk):
    """GroupbyCoalesce (the reference's CoalesceGroups layout, HostTasks.cpp:131-197) after the
    HIP groupby: the whole coalesced block and the per-group counts equal the layout built from
    the oracle's groups byte for byte."""
    from test_oracle_groupby import coalesce_brute, load_coalesce_dump, load_groupby_dump_prefix
    d = str(tmp_path / "c.bin")
    r = run_app("--kernels", paths.kernel_path("groupby.hsaco"), "--rows", n, "--num-keys", nk,
                "--dist", dist, "--coalesce", "--dump", d)
    assert r["ok"] and r["backend"] == "hip"
    keys, vals, groups, ok, ov = load_groupby_dump_prefix(d, n, nk)
    g, head, offsets, eok, eov = oracle.groupby(keys, vals, nk)
    _, co, meta = load_coalesce_dump(d, n, nk)
    want, wmeta = coalesce_brute(g, eov)
    G = g.shape[0]
    assert np.array_equal(meta[:G], wmeta)
    assert np.array_equal(co[:want.size], want)


@pytest.mark.parametrize("dist,n,nk", [("zipf", 300_001, 65535), ("uniform", 5000, 99)])
def test_hip_coalesce_exact_sizes(tmp_path, dist, n, nk):
    """--coalesce-exact on the GPU: the coalesced block holds exactly rows + G words and the
    counts block G words (sizes from GroupbyCoalesceSizes through MF_ALLOCATION_SIZE meta
    ports); both blocks equal the layout built from the oracle's groups byte for byte."""
    from test_oracle_groupby import coalesce_brute, load_dump_by_sizes
    d = str(tmp_path / "e.bin")
    r = run_app("--kernels", paths.kernel_path("groupby.hsaco"), "--rows", n, "--num-keys", nk,
                "--dist", dist, "--coalesce-exact", "--dump", d)
    assert r["ok"] and r["backend"] == "hip"
    keys, vals, out = load_dump_by_sizes(d, n, r)
    g, head, offsets, eok, eov = oracle.groupby(keys, vals, nk)
    want, wmeta = coalesce_brute(g, eov)
    G = g.shape[0]
    co, meta = out["coalesced"].view(np.int32), out["groupCounts"].view(np.int32)
    assert co.size == int(head[1]) + G and meta.size == G
    assert np.array_equal(meta, wmeta)
    assert np.array_equal(co, want)
