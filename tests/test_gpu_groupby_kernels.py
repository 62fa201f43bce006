"""The five groupby task kernels launched directly through the C ABI, against the oracle, over
key ranges, skews, row counts and launch grids (including grids smaller than the work, so blocks
stride over chunks and buckets).

Bar: bit-exact for every integer output -- per-key int64 count and sum, the groups header and
records, offsets, the reshuffled keys and values (rows of a group in input order).  Merge and
Reshuffle take blocks of exactly 256 threads; the accumulate block size varies."""
import numpy as np
import pytest

import oracle
from ptask_amd import groupby as GB
from ptask_amd import paths

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gmod(hip):
    return hip.Module(paths.kernel_path("groupby.hsaco"))


def make_keys(kind, n, nk, seed):
    rng = np.random.default_rng(seed)
    K = nk + 1
    if kind == "zipf":
        p = np.arange(1, K + 1, dtype=np.float64) ** -1.1
        cdf = np.cumsum(p / p.sum())
        r = np.minimum(np.searchsorted(cdf, rng.random(n), side="right"), K - 1)
        keys = (r * 40503) % K
    elif kind == "one_key":
        keys = np.full(n, min(7, nk))
    elif kind == "invalid":
        keys = rng.integers(0, K, n)
        bad = rng.random(n) < 0.05
        keys[bad] = rng.choice([-1, -5, K, K + 3, 2**31 - 1, -2**31], bad.sum())
    else:
        keys = rng.integers(0, K, n)
    vals = rng.integers(-2**31, 2**31, n, dtype=np.int64)
    return keys.astype(np.int32), vals.astype(np.int32)


def dev(hip, a):
    b = hip.DeviceBuffer(max(a.nbytes, 16))
    if a.nbytes:
        b.upload(a)
    return b


def run_kernels(hip, gmod, keys, vals, nk, chunk_grid, merge_grid, off_grid, block):
    n = keys.size
    K = nk + 1
    NB = GB.buckets(nk)
    units = NB * GB.groups(n)
    table = dev(hip, np.full(3 * K + 1, 0x5A5A5A5A5A5A5A5A, np.int64))  # merge writes what it owns
    scratch = hip.DeviceBuffer(GB.scratch_bytes(n, nk))
    offsets = dev(hip, np.zeros(K, np.int64))
    groups = dev(hip, np.zeros(3 + 3 * K, np.int64))
    dk, dv = dev(hip, keys), dev(hip, vals)
    ok = dev(hip, np.zeros(n, np.int32))
    ov = dev(hip, np.zeros(n, np.int32))
    f = gmod.function
    hip.launch(f("GroupbyAccumulate"), (min(chunk_grid, GB.chunks(n)), 1, 1), (block, 1, 1),
               hip.ArgPack().ptr(dk.ptr).ptr(dv.ptr).i32(n).i32(nk).ptr(scratch.ptr))
    hip.launch(f("GroupbyMerge"), (min(merge_grid, units), 1, 1), (GB.BUCKET, 1, 1),
               hip.ArgPack().ptr(scratch.ptr).i32(n).i32(nk))
    hip.launch(f("GroupbyTotals"), (min(off_grid, NB), 1, 1), (GB.BUCKET, 1, 1),
               hip.ArgPack().ptr(scratch.ptr).i32(n).i32(nk).ptr(table.ptr))
    hip.launch(f("GroupbyOffsets"), (min(off_grid, NB), 1, 1), (GB.BUCKET, 1, 1),
               hip.ArgPack().ptr(table.ptr).i32(nk).ptr(scratch.ptr).ptr(offsets.ptr).ptr(groups.ptr))
    hip.launch(f("GroupbyReshuffle"), (min(chunk_grid, GB.chunks(n)), 1, 1), (GB.PLACE_THREADS, 1, 1),
               hip.ArgPack().ptr(dk.ptr).ptr(dv.ptr).i32(n).i32(nk).ptr(offsets.ptr).ptr(scratch.ptr)
               .ptr(ok.ptr).ptr(ov.ptr))
    hip.synchronize()
    out = {}
    for name, buf, a in (("table", table, np.zeros(3 * K + 1, np.int64)),
                         ("offsets", offsets, np.zeros(K, np.int64)),
                         ("groups", groups, np.zeros(3 + 3 * K, np.int64)),
                         ("ok", ok, np.zeros(n, np.int32)), ("ov", ov, np.zeros(n, np.int32))):
        buf.download(a)
        out[name] = a
    return out


# 1M rows span two merge groups of 256 chunks, 2.1M rows five (the last one short)
CASES = [("zipf", 200_003, 65535), ("zipf", 1_000_000, GB.MAX_KEYS - 1), ("uniform", 4097, 0),
         ("invalid", 10_000, 1000), ("one_key", 50_001, 300), ("uniform", 2049, 513),
         ("uniform", 3, 7), ("uniform", 300_000, 299_999), ("zipf", 2048, 255),
         ("zipf", 2_100_000, 65535), ("one_key", 1_050_000, 0)]
GRIDS = {"full": (2048, 1 << 20, 1 << 20, 512), "strided": (3, 2, 2, 256), "wide": (2048, 7, 5, 1024)}


@pytest.mark.parametrize("grid", sorted(GRIDS))
@pytest.mark.parametrize("kind,n,nk", CASES)
def test_groupby_kernels_equal_oracle(hip, gmod, kind, n, nk, grid):
    print(f"case groupby kernels {kind} n={n} nk={nk} grid={grid}", flush=True)
    keys, vals = make_keys(kind, n, nk, seed=n + nk)
    got = run_kernels(hip, gmod, keys, vals, nk, *GRIDS[grid])
    g, head, offsets, eok, eov = oracle.groupby(keys, vals, nk)
    K = nk + 1
    cnt = np.zeros(K, np.int64)
    sm = np.zeros(K, np.int64)
    cnt[g[:, 0]] = g[:, 1]
    sm[g[:, 0]] = g[:, 2]
    assert np.array_equal(got["table"][:K], cnt)
    assert np.array_equal(got["table"][K:2 * K], sm)
    assert got["table"][3 * K] == head[2]
    assert np.array_equal(got["groups"][:3], head)
    assert np.array_equal(got["groups"][3:3 + g.size], g.reshape(-1))
    assert np.array_equal(got["offsets"], offsets)
    rows = int(head[1])
    assert np.array_equal(got["ok"][:rows], eok[:rows])
    assert np.array_equal(got["ov"][:rows], eov[:rows])
