"""The groupby pipeline as direct library calls (include/ptk_groupby_api.h, libptk_groupby.so).

CPU part: the library loads, exports every declared entry point, sizes its workspace as the
table plus the scratch of include/ptk_groupby.h, and rejects bad arguments before any HIP call.
GPU part: one call on device buffers gives the oracle's groups, offsets and stable grouping bit
for bit, at several shapes, and equals the groupby graph's kernels run task by task.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from ptask_amd import groupby as GB


def test_library_exports_every_declared_symbol():
    names = GB.declared_symbols()
    assert set(names) >= {"ptk_groupby_workspace_bytes", "ptk_groupby_run", "ptk_groupby_last_error"}
    L = GB.lib()
    for n in names:
        assert hasattr(L, n), n


def test_python_layout_constants_equal_the_header():
    """ptask_amd/groupby.py mirrors the PTK_GB_* constants of include/ptk_groupby.h."""
    import os
    import re

    from ptask_amd import paths
    text = open(os.path.join(paths.include_dir(), "ptk_groupby.h")).read()
    macros = dict(re.findall(r"^#define PTK_GB_(\w+) (\d+)", text, re.M))
    for name in ("CHUNK", "SUB", "TILE", "BUCKET", "HIST_THREADS", "COUNT_THREADS", "PART_THREADS",
                 "TILE_THREADS", "SCAN_THREADS", "KEYSCAN_THREADS", "TOTALS_THREADS", "MAX_BUCKETS"):
        assert int(macros[name]) == getattr(GB, name), name


@pytest.mark.parametrize("n,nk", [(1, 0), (1000, 99), (2_100_000, 65535), (300_000, GB.MAX_KEYS - 2)])
def test_workspace_is_table_plus_scratch(n, nk):
    up = lambda x: (x + 255) // 256 * 256
    assert GB.lib().ptk_groupby_workspace_bytes(n, nk) == up(8 * (3 * (nk + 1) + 1)) + GB.scratch_bytes(n, nk)


def test_bad_arguments_are_rejected_before_any_device_call():
    L = GB.lib()
    assert L.ptk_groupby_workspace_bytes(0, 5) == -1
    assert L.ptk_groupby_workspace_bytes(10, GB.MAX_KEYS) == -1
    p = C.c_void_p(0x1000)  # never dereferenced: every check below fails first
    assert L.ptk_groupby_run(p, p, 0, 5, p, 1 << 30, p, p, p, p, None) == GB.EINVAL
    assert L.ptk_groupby_run(p, p, 10, -1, p, 1 << 30, p, p, p, p, None) == GB.EINVAL
    assert L.ptk_groupby_run(None, p, 10, 5, p, 1 << 30, p, p, p, p, None) == GB.EINVAL
    assert L.ptk_groupby_run(p, p, 10, 5, C.c_void_p(0x1008), 1 << 30, p, p, p, p, None) == GB.EINVAL
    assert b"aligned" in L.ptk_groupby_last_error()
    need = L.ptk_groupby_workspace_bytes(10, 5)
    assert L.ptk_groupby_run(p, p, 10, 5, p, need - 1, p, p, p, p, None) == GB.ESPACE
    assert b"needed" in L.ptk_groupby_last_error()


def zipf_rows(n, nk, seed):
    rng = np.random.default_rng(seed)
    K = nk + 1
    p = np.arange(1, K + 1, dtype=np.float64) ** -1.1
    r = np.minimum(np.searchsorted(np.cumsum(p / p.sum()), rng.random(n), side="right"), K - 1)
    keys = ((r * 40503) % K).astype(np.int32)
    bad = rng.random(n) < 0.01
    keys[bad] = -3
    vals = rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)
    return keys, vals


def run_direct(hip, keys, vals, nk, fn="run"):
    """One library call on device buffers; returns (groups, offsets, outKeys, outVals)."""
    n, K = keys.size, nk + 1
    L = GB.lib()
    dk, dv = hip.DeviceBuffer(keys.nbytes), hip.DeviceBuffer(vals.nbytes)
    dk.upload(keys)
    dv.upload(vals)
    off, grp = hip.DeviceBuffer(8 * K), hip.DeviceBuffer(8 * (3 + 3 * K))
    ok, ov = hip.DeviceBuffer(4 * n), hip.DeviceBuffer(4 * n)
    if fn == "run":
        ws_bytes = L.ptk_groupby_workspace_bytes(n, nk)
        ws = hip.DeviceBuffer(ws_bytes)
        rc = L.ptk_groupby_run(dk.ptr, dv.ptr, n, nk, ws.ptr, ws_bytes, off.ptr, grp.ptr, ok.ptr, ov.ptr, None)
    else:
        rc = L.ptk_groupby_run_noalloc(dk.ptr, dv.ptr, n, nk, off.ptr, grp.ptr, ok.ptr, ov.ptr, None)
    assert rc == 0, L.ptk_groupby_last_error()
    hip.synchronize()
    out = [np.zeros(3 + 3 * K, np.int64), np.zeros(K, np.int64), np.zeros(n, np.int32), np.zeros(n, np.int32)]
    for b, a in zip((grp, off, ok, ov), out):
        b.download(a)
    return out


def check_against_oracle(keys, vals, nk, got):
    groups, offsets, ok, ov = got
    g, head, eoff, eok, eov = oracle.groupby(keys, vals, nk)
    rows = int(head[1])
    assert np.array_equal(groups[:3], head)
    assert np.array_equal(groups[3:3 + g.size], g.reshape(-1))
    assert np.array_equal(offsets, eoff)
    assert np.array_equal(ok[:rows], eok[:rows]) and np.array_equal(ov[:rows], eov[:rows])


@pytest.mark.gpu
@pytest.mark.parametrize("n,nk", [(1, 0), (5000, 99), (1_000_003, 65535), (2_100_000, 4095)])
def test_direct_call_equals_oracle(hip, n, nk):
    print(f"case groupby direct n={n} nk={nk}", flush=True)
    keys, vals = zipf_rows(n, nk, seed=n)
    check_against_oracle(keys, vals, nk, run_direct(hip, keys, vals, nk))


def test_noalloc_rejects_inputs_beyond_its_static_workspace():
    L = GB.lib()
    cap = L.ptk_groupby_noalloc_capacity()
    assert cap == 64 << 20
    p = C.c_void_p(0x1000)  # never dereferenced: the size check fails first
    n = 12_000_000
    assert L.ptk_groupby_workspace_bytes(n, 65535) > cap
    assert L.ptk_groupby_run_noalloc(p, p, n, 65535, p, p, p, p, None) == GB.ESPACE
    assert b"static workspace" in L.ptk_groupby_last_error()
    assert L.ptk_groupby_run_noalloc(p, p, 0, 65535, p, p, p, p, None) == GB.EINVAL
    assert L.ptk_groupby_run_noalloc(None, p, 10, 5, p, p, p, p, None) == GB.EINVAL
