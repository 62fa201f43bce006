"""The reference's groupby graph (ptaskUnitTest/graphgroupby.cpp:56-573) through the PTask runtime
(ptask_amd/csrc/apps/ptk_groupby_ref.cpp): the same seven tasks, ports, channels, priorities and
affinity, bound by name to countUniqueKeysIntInt ... deallocateCountArray.

The reference checks only that equal keys are contiguous and the key set is kept
(graphgroupby.cpp:394-419); the app applies that check and reports it.  The groups here come in
ascending key order, so the output must equal numpy's sort of the input bit for bit: on the host
through the oracle's host-task forms, on the GPU through groupby_ref.hsaco.  Inputs: the
reference's own configuration (1024 rows, 32 keys, MSVC rand() seeded 0x4fd478: regtest.sh:1115,
graphgroupby.cpp:428,555), keys spread over the whole int32 range with negatives (the radix-sort
ranking), every row distinct, and repeated runs through one graph (device heap reuse).
"""
import json
import subprocess

import numpy as np
import pytest

from ptask_amd import paths
from test_pca_graph import HOSTLIB

APP = paths.bin_path("ptk_groupby_ref")


def run_ref(kernels, tmp_path, rows, keys, iters, dist, timeout=300):
    dump = str(tmp_path / f"gbr_{rows}_{keys}_{dist}.bin")
    cmd = [APP, "--kernels", kernels, "-r", str(rows), "-c", str(keys), "-i", str(iters), "--dist", dist,
           "--dump", dump]
    print("run:", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    assert p.returncode == 0, p.stderr + p.stdout
    js = json.loads(p.stdout.strip().splitlines()[-1])
    d = np.fromfile(dump, dtype=np.int32)
    assert d.size == 2 * rows
    return js, d[:rows], d[rows:]


def check(js, keys_in, out):
    assert js["ok"] and js["succeeded"] and js["failures"] == 0 and js["same_key_set"]
    want = np.sort(keys_in, kind="stable")
    bad = np.flatnonzero(out != want)
    assert bad.size == 0, f"{bad.size} rows differ from the sorted input, first at {bad[:4]}"


def test_reference_keys_match_msvc_rand(tmp_path):
    # the reference's inputs: (int)(rand()/(float)RAND_MAX * numKeys), MSVC rand() seeded 0x4fd478
    holdrand, want = 0x4fd478, []
    for _ in range(8):
        holdrand = (holdrand * 214013 + 2531011) & 0xFFFFFFFF
        r = (holdrand >> 16) & 0x7FFF
        want.append(int(np.float32(r) / np.float32(32767) * np.float32(32)))
    js, keys_in, out = run_ref(HOSTLIB, tmp_path, 8, 32, 1, "reference")
    assert keys_in.tolist() == want


@pytest.mark.parametrize("rows,keys,iters,dist", [(1024, 32, 3, "reference"), (5000, 77, 2, "wide"),
                                                  (3000, 1, 1, "distinct"), (1, 5, 1, "reference")])
def test_host_graph(tmp_path, rows, keys, iters, dist):
    js, keys_in, out = run_ref(HOSTLIB, tmp_path, rows, keys, iters, dist)
    assert js["backend"] == "host"
    check(js, keys_in, out)


def test_usage_errors():
    p = subprocess.run([APP, "--kernels", HOSTLIB, "--dist", "zipf"], capture_output=True, text=True)
    assert p.returncode == 2 and "usage" in p.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("rows,keys,iters,dist", [
    (1024, 32, 3, "reference"),          # the reference's configuration
    (1_000_000, 65536, 2, "reference"),  # dense key range: bitmap ranking
    (200_000, 50_000, 2, "wide"),        # int32-wide keys with negatives: radix-sort ranking
    (100_000, 1, 1, "distinct"),         # every row its own key
    (2_000_000, 1000, 6, "reference"),   # six runs through one graph: the device heap is reused
])
def test_gpu_graph(tmp_path, rows, keys, iters, dist):
    print(f"case groupby_ref rows={rows} keys={keys} iters={iters} dist={dist}", flush=True)
    js, keys_in, out = run_ref(paths.kernel_path("groupby_ref.hsaco"), tmp_path, rows, keys, iters, dist, timeout=600)
    assert js["backend"] == "hip"
    check(js, keys_in, out)
