"""The seven groupby task kernels launched directly through the C ABI, against the oracle, over
key ranges (up to 2^19 keys: 2048 buckets), skews, row counts and launch grids (including grids
smaller than the work, so blocks stride over chunks, buckets and tiles).

Bar: bit-exact for every integer output -- per-key int64 count and sum, the groups header and
records, offsets, the reshuffled keys and values (rows of a group in input order)."""
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


def run_kernels(hip, gmod, keys, vals, nk, grids):
    """The seven launches in the graph's order; grids[i] caps launch i's grid (None: full)."""
    n = keys.size
    K = nk + 1
    NB = GB.buckets(nk)
    TM = GB.tiles_max(n, nk)
    table = dev(hip, np.full(3 * K + 1, 0x5A5A5A5A5A5A5A5A, np.int64))  # totals writes what it owns
    scratch = hip.DeviceBuffer(GB.scratch_bytes(n, nk))
    offsets = dev(hip, np.zeros(K, np.int64))
    groups = dev(hip, np.zeros(3 + 3 * K, np.int64))
    dk, dv = dev(hip, keys), dev(hip, vals)
    ok = dev(hip, np.zeros(n, np.int32))
    ov = dev(hip, np.zeros(n, np.int32))
    f = gmod.function
    full = [max(1, GB.chunks(n)), NB, max(1, min(GB.chunks(n), 512)), TM, 2 * NB, TM, 4 * NB]
    g = [full[i] if grids[i] is None else max(1, min(grids[i], full[i])) for i in range(7)]
    sc = lambda: hip.ArgPack().ptr(scratch.ptr).i32(n).i32(nk)
    hip.launch(f("GroupbyHistogram"), (g[0], 1, 1), (GB.HIST_THREADS, 1, 1),
               hip.ArgPack().ptr(dk.ptr).i32(n).i32(nk).ptr(scratch.ptr))
    hip.launch(f("GroupbyBucketScan"), (g[1], 1, 1), (GB.SCAN_THREADS, 1, 1), sc())
    hip.launch(f("GroupbyPartition"), (g[2], 1, 1), (GB.PART_THREADS, 1, 1),
               hip.ArgPack().ptr(dk.ptr).ptr(dv.ptr).i32(n).i32(nk).ptr(scratch.ptr))
    hip.launch(f("GroupbyKeyCount"), (g[3], 1, 1), (GB.COUNT_THREADS, 1, 1), sc())
    hip.launch(f("GroupbyKeyScan"), (g[4], 1, 1), (GB.KEYSCAN_THREADS, 1, 1), sc().ptr(offsets.ptr))
    hip.launch(f("GroupbyReshuffle"), (g[5], 1, 1), (GB.TILE_THREADS, 1, 1),
               sc().ptr(offsets.ptr).ptr(ok.ptr).ptr(ov.ptr))
    hip.launch(f("GroupbyTotals"), (g[6], 1, 1), (GB.TOTALS_THREADS, 1, 1), sc().ptr(table.ptr).ptr(groups.ptr))
    hip.synchronize()
    out = {}
    for name, buf, a in (("table", table, np.zeros(3 * K + 1, np.int64)),
                         ("offsets", offsets, np.zeros(K, np.int64)),
                         ("groups", groups, np.zeros(3 + 3 * K, np.int64)),
                         ("ok", ok, np.zeros(n, np.int32)), ("ov", ov, np.zeros(n, np.int32))):
        buf.download(a)
        out[name] = a
    return out


# 2.1M rows: 129 chunks of 16384 rows (the last one short), hundreds of tiles; one_key puts a whole
# input into one bucket (hundreds of tiles of one key); 2^19 keys use 2048 buckets
CASES = [("zipf", 200_003, 65535), ("zipf", 1_000_000, GB.MAX_KEYS - 1), ("uniform", 4097, 0),
         ("invalid", 10_000, 1000), ("one_key", 50_001, 300), ("uniform", 2049, 513),
         ("uniform", 3, 7), ("uniform", 300_000, 299_999), ("zipf", 2048, 255),
         ("zipf", 2_100_000, 65535), ("one_key", 1_050_000, 0), ("invalid", 70_000, GB.MAX_KEYS - 2)]
GRIDS = {"full": [None] * 7, "strided": [3, 2, 1, 3, 2, 5, 3], "wide": [None, None, 7, None, 5, 11, 6]}


@pytest.mark.parametrize("grid", sorted(GRIDS))
@pytest.mark.parametrize("kind,n,nk", CASES)
def test_groupby_kernels_equal_oracle(hip, gmod, kind, n, nk, grid):
    print(f"case groupby kernels {kind} n={n} nk={nk} grid={grid}", flush=True)
    keys, vals = make_keys(kind, n, nk, seed=n + nk)
    got = run_kernels(hip, gmod, keys, vals, nk, GRIDS[grid])
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
