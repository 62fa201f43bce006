"""Pins the prefix-sum and groupby oracle (oracle/groupby.c), then checks the groupby and
prefix-sum graphs on the HostAccelerator against it.

Prefix sum: the reference's own fixtures (g_data of pfxsum.cpp:22-40, g_cdata of
cpfxsum.cpp:25-154, extracted by tests/golden/make_pfxsum_fixtures.py) with their closed-form
exclusive scans, an independent numpy cumsum, and int32 wrap-around.
Groupby: an independent numpy formulation (np.unique counts, np.add.at sums, a stable argsort
for the reshuffled layout) and the reference's own acceptance invariant (graphgroupby.cpp:394-419:
equal keys contiguous, same key set).
"""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from ptask_amd import paths

GOLD = os.path.join(os.path.dirname(__file__), "golden", "pfxsum_fixtures.json")
APP = paths.bin_path("ptk_groupby")
HOSTLIB = os.path.join(oracle.ORACLE_DIR, "_build", "libptk_oracle.so")


def test_pfxsum_reference_fixtures():
    fx = json.load(open(GOLD))
    for name in ("g_data", "g_cdata"):
        a = np.array(fx[name]["input"], dtype=np.int32)
        got = oracle.pfxsum(a)
        assert got.tolist() == fx[name]["exclusive_scan"]
        # closed form: out[i] = sum_{j<i} in[j]
        assert got.tolist() == [int(a[:i].sum()) for i in range(a.size)]
    used = fx["g_cdata"]["used_prefix"]
    a = np.array(fx["g_cdata"]["input"][:used], dtype=np.int32)
    assert oracle.pfxsum(a).tolist() == fx["g_cdata"]["exclusive_scan"][:used]


def test_pfxsum_wraps_like_int32():
    a = np.array([2**31 - 1, 1, 5, -7, 2**31 - 1], dtype=np.int32)
    got = oracle.pfxsum(a)
    ref = np.concatenate([[0], np.cumsum(a.astype(np.int64))[:-1]])
    assert got.tolist() == [int(((x + 2**31) % 2**32) - 2**31) for x in ref]
    b = np.arange(1, 200000, dtype=np.int64) * 7919
    assert np.array_equal(oracle.pfxsum(b), np.concatenate([[0], np.cumsum(b)[:-1]]))
    assert oracle.pfxsum(np.zeros(0, np.int32)).size == 0


def reference_invariant(inp, out):
    """testResults (graphgroupby.cpp:394-419): contiguity + same key set."""
    seen, failures = set(), 0
    for i in range(len(out) - 1):
        if out[i] in seen:
            failures += 1
        if out[i] != out[i + 1]:
            seen.add(out[i])
    if len(out):
        if out[-1] in seen:
            failures += 1
        seen.add(out[-1])
    return failures == 0 and seen == set(inp.tolist())


@pytest.mark.parametrize("n,nk", [(1, 0), (17, 3), (1024, 32), (50000, 999), (20000, 65535)])
def test_groupby_matches_numpy_formulation(n, nk):
    rng = np.random.default_rng(n + nk)
    keys = rng.integers(0, nk + 1, n).astype(np.int32)
    vals = rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)
    groups, head, offsets, ok, ov = oracle.groupby(keys, vals, nk)
    uk, cnt = np.unique(keys, return_counts=True)
    sums = np.zeros(nk + 1, dtype=np.int64)
    np.add.at(sums, keys, vals.astype(np.int64))
    assert head.tolist() == [uk.size, n, 0]
    assert np.array_equal(groups[:, 0], uk) and np.array_equal(groups[:, 1], cnt)
    assert np.array_equal(groups[:, 2], sums[uk])
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(ok, keys[order]) and np.array_equal(ov, vals[order])
    full = np.zeros(nk + 1, dtype=np.int64)
    full[uk] = cnt
    assert np.array_equal(offsets, np.concatenate([[0], np.cumsum(full)[:-1]]))
    assert reference_invariant(keys, ok)


def test_groupby_counts_out_of_range_keys_as_errors():
    keys = np.array([0, 5, -1, 3, 6, 3], dtype=np.int32)
    vals = np.array([1, 2, 3, 4, 5, 6], dtype=np.int32)
    groups, head, offsets, ok, ov = oracle.groupby(keys, vals, 5)
    assert head.tolist() == [3, 4, 2]
    assert groups.tolist() == [[0, 1, 1], [3, 2, 10], [5, 1, 2]]


def run_app(*args, timeout=300):
    p = subprocess.run([APP, *map(str, args)], capture_output=True, text=True, timeout=timeout)
    assert p.returncode == 0, p.stderr + p.stdout
    return json.loads(p.stdout.strip().splitlines()[-1])


def load_groupby_dump(path, n, nk):
    raw = open(path, "rb").read()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a
    keys, vals = take(np.int32, n), take(np.int32, n)
    groups = take(np.int64, 3 + 3 * (nk + 1))
    ok, ov = take(np.int32, n), take(np.int32, n)
    assert off == len(raw)
    return keys, vals, groups, ok, ov


@pytest.mark.parametrize("dist", ["zipf", "uniform"])
def test_host_groupby_graph_equals_oracle(tmp_path, dist):
    n, nk = 30000, 1023
    d = str(tmp_path / "g.bin")
    r = run_app("--kernels", HOSTLIB, "--rows", n, "--num-keys", nk, "--dist", dist, "--repeat", 2,
                "--dump", d)
    assert r["ok"] and r["backend"] == "host"
    keys, vals, groups, ok, ov = load_groupby_dump(d, n, nk)
    g, head, offsets, eok, eov = oracle.groupby(keys, vals, nk)
    assert np.array_equal(groups[:3], head)
    assert np.array_equal(groups[3:3 + g.size], g.reshape(-1))
    assert np.array_equal(ok, eok) and np.array_equal(ov, eov)
    if dist == "zipf":  # skewed: the most frequent key holds a large share of the rows
        assert g[:, 1].max() > 0.05 * n


def test_host_pfxsum_graph_equals_oracle(tmp_path):
    for n in (24, 128, 70001):
        d = str(tmp_path / f"p{n}.bin")
        r = run_app("--kernels", HOSTLIB, "--workload", "pfxsum", "--rows", n, "--dump", d)
        raw = np.fromfile(d, dtype=np.int32)
        keys, out = raw[:n], raw[2 * n:3 * n]
        assert np.array_equal(out, oracle.pfxsum(keys)), n


@pytest.mark.parametrize("n,nk", [(1, 0), (2049, 513), (10_000, 1000), (6000, 255), (1_100_000, 4095),
                                  (70_000, 2**19 - 2)])
def test_host_task_forms_pipeline(n, nk):
    """The host-task forms of the seven groupby tasks (the two-level radix partition and scratch
    layout of include/ptk_groupby.h restated on the host) chained as the graph chains them, with
    invalid keys, equal the oracle's pipeline exactly, values in input order included."""
    import ctypes as C

    from ptask_amd import groupby as GB
    L = C.CDLL(HOSTLIB)
    rng = np.random.default_rng(n)
    K = nk + 1
    keys = rng.integers(0, K, n).astype(np.int32)
    bad = rng.random(n) < 0.05
    keys[bad] = rng.choice([-1, K, 2**31 - 1], bad.sum()).astype(np.int32)
    vals = rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)
    table = np.full(3 * K + 1, 0x5A5A5A5A5A5A5A5A, np.int64)  # totals writes what it owns
    scratch = np.zeros(GB.scratch_bytes(n, nk) // 8 + 2, np.int64)
    offsets = np.zeros(K, np.int64)
    groups = np.zeros(3 + 3 * K, np.int64)
    ok, ov = np.zeros(n, np.int32), np.zeros(n, np.int32)

    def call(name, *args):
        arr = (C.c_void_p * len(args))(*[C.c_void_p(a) for a in args])
        getattr(L, name)(C.c_uint(len(args)), arr)

    p = lambda a: a.ctypes.data
    call("GroupbyHistogram", p(keys), n, nk, p(scratch))
    call("GroupbyBucketScan", p(scratch), n, nk)
    call("GroupbyPartition", p(keys), p(vals), n, nk, p(scratch))
    call("GroupbyKeyCount", p(scratch), n, nk)
    call("GroupbyKeyScan", p(scratch), n, nk, p(offsets))
    call("GroupbyReshuffle", p(scratch), n, nk, p(offsets), p(ok), p(ov))
    call("GroupbyTotals", p(scratch), n, nk, p(table), p(groups))
    g, head, eoff, eok, eov = oracle.groupby(keys, vals, nk)
    assert np.array_equal(groups[:3], head) and head[2] == bad.sum()
    assert np.array_equal(groups[3:3 + g.size], g.reshape(-1))
    assert np.array_equal(offsets, eoff)
    rows = int(head[1])
    assert np.array_equal(ok[:rows], eok[:rows]) and np.array_equal(ov[:rows], eov[:rows])


def coalesce_brute(g, eov):
    """The coalesced layout built directly: per group its key, then its values."""
    words, pos = [], 0
    for key, cnt, _ in g:
        words.append(np.int32(key))
        words.extend(eov[pos:pos + cnt].view(np.int32))
        pos += cnt
    return np.array(words, dtype=np.int32), g[:, 1].astype(np.int32)


def load_coalesce_dump(path, n, nk):
    raw = np.fromfile(path, dtype=np.uint8)
    off = 4 * n * 2 + 8 * (3 + 3 * (nk + 1)) + 4 * n * 2
    head = raw[:off]
    co = raw[off:off + 4 * (n + nk + 1)].view(np.int32)
    meta = raw[off + 4 * (n + nk + 1):].view(np.int32)
    return head, co, meta


def test_coalesce_int_float_restates_the_reference_layout():
    """coalesce_int_float (HostTasks.cpp:164-197 ABI) writes key, values per group, groups
    consecutive, and the per-group counts; pinned against the layout built directly."""
    import ctypes as C
    L = C.CDLL(HOSTLIB)
    ids = np.array([4, 4, 4, 7, 9, 9], np.int32)
    vals = np.array([1.5, -2.0, 3.25, 7.0, 0.5, 1e30], np.float32)
    counts = np.array([3, 1, 2], np.int32)
    out = np.full(9, -1, np.int32)
    meta = np.zeros(3, np.int32)
    args = (C.c_void_p * 7)(vals.ctypes.data, 6, ids.ctypes.data, 3, counts.ctypes.data, out.ctypes.data,
                            meta.ctypes.data)
    L.coalesce_int_float(C.c_uint(7), args)
    want = np.concatenate([[4], vals[:3].view(np.int32), [7], vals[3:4].view(np.int32), [9], vals[4:].view(np.int32)])
    assert np.array_equal(out, want.astype(np.int32)) and np.array_equal(meta, counts)


def test_host_groupby_graph_coalesce(tmp_path):
    n, nk = 20000, 511
    d = str(tmp_path / "c.bin")
    r = run_app("--kernels", HOSTLIB, "--rows", n, "--num-keys", nk, "--coalesce", "--dump", d)
    assert r["ok"]
    keys, vals, groups, ok, ov = load_groupby_dump_prefix(d, n, nk)
    g, head, offsets, eok, eov = oracle.groupby(keys, vals, nk)
    _, co, meta = load_coalesce_dump(d, n, nk)
    want, wmeta = coalesce_brute(g, eov)
    G = g.shape[0]
    assert np.array_equal(co[:want.size], want) and np.array_equal(meta[:G], wmeta)


def load_dump_by_sizes(path, n, r):
    """Inputs (keys, vals) and each output of a groupby run, cut by the sizes the run reports."""
    raw = np.fromfile(path, dtype=np.uint8)
    keys, vals = raw[:4 * n].view(np.int32), raw[4 * n:8 * n].view(np.int32)
    out, off = {}, 8 * n
    for o in r["outputs"]:
        out[o["name"]] = raw[off:off + o["bytes"]]
        off += o["bytes"]
    assert off == raw.size
    return keys, vals, out


@pytest.mark.parametrize("n,nk", [(20000, 511), (777, 0), (5000, 65535)])
def test_host_groupby_graph_coalesce_exact_sizes(tmp_path, n, nk):
    """--coalesce-exact: GroupbyCoalesceSizes feeds MF_ALLOCATION_SIZE meta ports, so the
    coalesced block holds exactly rows + G words and the counts block G words (the reference's
    variable-size CoalesceGroups output), with the same content as the fixed-capacity form."""
    d = str(tmp_path / "e.bin")
    r = run_app("--kernels", HOSTLIB, "--rows", n, "--num-keys", nk, "--coalesce-exact", "--dump", d)
    assert r["ok"]
    keys, vals, out = load_dump_by_sizes(d, n, r)
    g, head, offsets, eok, eov = oracle.groupby(keys, vals, nk)
    want, wmeta = coalesce_brute(g, eov)
    G = g.shape[0]
    assert out["coalesced"].size == 4 * (int(head[1]) + G) and out["groupCounts"].size == 4 * G
    assert np.array_equal(out["coalesced"].view(np.int32), want)
    assert np.array_equal(out["groupCounts"].view(np.int32), wmeta)


def load_groupby_dump_prefix(path, n, nk):
    raw = np.fromfile(path, dtype=np.uint8)
    off = 0

    def take(dtype, count):
        nonlocal off
        a = raw[off:off + count * np.dtype(dtype).itemsize].view(dtype)
        off += a.nbytes
        return a
    keys, vals = take(np.int32, n), take(np.int32, n)
    groups = take(np.int64, 3 + 3 * (nk + 1))
    return keys, vals, groups, take(np.int32, n), take(np.int32, n)
