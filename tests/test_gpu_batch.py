"""The batched / served executor (include/ptk_batch.h) on the GPU.

Bar: every instance computes what its graph computes alone -- FDTD fields bit-exact against the
oracle, every groupby output bit-exact (values of a group in input order), GMM within the tolerance of the single-graph GMM kernels (log-likelihood 1e-4
relative) -- in plain per-task dispatch and in batched captured iterations alike, and the
result checksums of an instance are identical alone, in a mix, in either mode, and when
admission by footprint delays it.
"""
import numpy as np
import pytest

import oracle
from prng import rand_unit_f64
from ptask_amd import batch, gmm
from test_gpu_gmm import check_against_oracle

pytestmark = pytest.mark.gpu

GB = 1 << 30
MODES = [batch.PLAIN, batch.BATCHED]


@pytest.mark.parametrize("tiles,launch_steps", [((64, 4), 1), ((0, 3), 1), ((0, 1), 1), ((0, 64), 1),
                                                ((0, 3), 3), ((0, 8), 3)])
@pytest.mark.parametrize("mode", MODES)
def test_fdtd_instances_bit_exact(hip, mode, tiles, launch_steps):
    """Half-step column tiles (cols > 0) and whole-step overlapped tiles (cols == 0) of several
    depths, one or three steps per launch (three while every instance has three left, then single
    steps; step counts that are not multiples of three); small tiles put several per instance."""
    print(f"case batch fdtd mode={mode} tiles={tiles} launch_steps={launch_steps}", flush=True)
    shapes = [(12, 10, 8, 5, 11), (16, 16, 16, 4, 12), (20, 17, 9, 3, 13), (5, 33, 7, 6, 14), (70, 13, 6, 2, 15),
              (61, 23, 19, 10, 16), (1, 1, 1, 7, 17)]
    with batch.Executor(1 * GB, mode) as ex:
        ex.set_tiles(*tiles, 65536)
        ex.set_fdtd_steps(launch_steps)
        ids = [ex.submit_fdtd(nx, ny, nz, s, seed, keep=True) for nx, ny, nz, s, seed in shapes]
        ex.drain()
        for i, (nx, ny, nz, s, seed) in zip(ids, shapes):
            want = oracle.fdtd_run(oracle.init_fields(nx, ny, nz, seed), oracle.make_params(nx, ny, nz, 0.5), s)
            for f in range(6):
                assert np.array_equal(ex.read(i, f, np.float32), want[f]), (nx, ny, nz, f)


@pytest.mark.parametrize("mode", MODES)
def test_groupby_instances_match_oracle(hip, mode):
    print(f"case batch groupby mode={mode}", flush=True)
    cases = [(100_000, 4095, 1.1, 21), (50_001, 999, 0.0, 22), (300_000, 65535, 1.1, 23)]
    with batch.Executor(1 * GB, mode) as ex:
        ex.set_tiles(512, 16, 32768)
        ids = [ex.submit_groupby(r, nk, s, seed, keep=True) for r, nk, s, seed in cases]
        ex.drain()
        for i, (rows, nk, s, seed) in zip(ids, cases):
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
            c = ex.checksums(i)
            assert c[4] == c[3]  # the output rows are a permutation of the input rows


def gmm_initial_means(X, K, seed):
    """Means of the executor's initial model: the first point of each generating component."""
    N = X.shape[0]
    comp = np.minimum((rand_unit_f64(seed, 12, np.arange(N)) * K).astype(np.int64), K - 1)
    mu0 = np.empty((K, X.shape[1]))
    for c in range(K):
        idx = np.flatnonzero(comp == c)
        mu0[c] = X[idx[0] if idx.size else (c * N) // K]
    return mu0


@pytest.mark.parametrize("mode", MODES)
def test_gmm_instances_match_oracle(hip, mode):
    print(f"case batch gmm mode={mode}", flush=True)
    cases = [(20_000, 8, 6, 6, 31), (4_097, 16, 12, 5, 32)]
    with batch.Executor(1 * GB, mode) as ex:
        ids = [ex.submit_gmm(n, d, k, s, seed, keep=True) for n, d, k, s, seed in cases]
        ex.drain()
        for i, (n, d, k, s, seed) in zip(ids, cases):
            X = ex.read(i, 0, np.float32).reshape(n, d)
            m = gmm.parse_model(ex.read(i, 1, np.float64), d, k, s)
            check_against_oracle(X, gmm_initial_means(X, k, seed), m, s, f"batch gmm n={n} d={d} k={k}")


MIX = [(batch.FDTD, (24, 24, 24, 7)), (batch.GROUPBY, (200_000, 65535, 1.1)), (batch.FDTD, (40, 31, 17, 3)),
       (batch.GMM, (30_000, 16, 8, 4)), (batch.FDTD, (16, 16, 16, 9)), (batch.GROUPBY, (77_777, 4095, 0.0)),
       (batch.GMM, (12_345, 16, 32, 6)), (batch.FDTD, (33, 9, 50, 5))]


def submit(ex, kind, args, seed, keep=False):
    if kind == batch.FDTD:
        return ex.submit_fdtd(*args, seed, keep)
    if kind == batch.GROUPBY:
        return ex.submit_groupby(*args, seed, keep)
    return ex.submit_gmm(*args, seed, keep)


def run_mix(mode, arena, tiles=None, autotune=False, graphs=True, launch_steps=3):
    with batch.Executor(arena, mode) as ex:
        ex.set_launch(graphs)
        ex.set_fdtd_steps(launch_steps)
        if tiles:
            ex.set_tiles(*tiles)
        ids = [submit(ex, kind, args, 100 + j) for j, (kind, args) in enumerate(MIX)]
        ex.tick()
        if autotune:
            ex.autotune()
        ex.drain()
        return [ex.checksums(i) for i in ids], ex.stats()


def test_checksums_equal_alone_in_mix_both_modes_and_delayed(hip):
    print("case batch identity", flush=True)
    mixed, st = run_mix(batch.BATCHED, 2 * GB)
    assert st["graph_launches"] >= 1
    plain, _ = run_mix(batch.PLAIN, 2 * GB)
    assert plain == mixed
    tuned, _ = run_mix(batch.BATCHED, 2 * GB, tiles=(128, 4, 16384), autotune=True)
    assert tuned == mixed
    halves, _ = run_mix(batch.BATCHED, 2 * GB, tiles=(256, 8, 262144))
    assert halves == mixed  # the half-step and whole-step FDTD forms agree bit for bit
    single, _ = run_mix(batch.BATCHED, 2 * GB, launch_steps=1)
    assert single == mixed  # one step per launch and three steps per launch agree bit for bit
    direct, st = run_mix(batch.BATCHED, 2 * GB, graphs=False)
    assert direct == mixed and st["graph_launches"] == 0
    # an arena that holds only a few instances at a time: admission by footprint queues the rest
    def fp(kind, a):
        if kind == batch.FDTD:
            return batch.footprint(kind, a[0], a[1], a[2], a[3])
        if kind == batch.GROUPBY:
            return batch.footprint(kind, a[0], a[1])
        return batch.footprint(kind, a[0], a[1], a[2], a[3])
    big = max(fp(k, a) for k, a in MIX)
    delayed, st = run_mix(batch.BATCHED, int(2.2 * big))
    assert delayed == mixed and st["arena_peak"] <= int(2.2 * big)
    for j, (kind, args) in enumerate(MIX):
        with batch.Executor(2 * GB, batch.BATCHED) as ex:
            i = submit(ex, kind, args, 100 + j)
            ex.drain()
            assert ex.checksums(i) == mixed[j], (kind, args)


def test_staggered_arrivals_reuse_memory_and_keep_results(hip):
    """Instances arriving while others run (continuous batching) in an arena small enough that
    later instances reuse the memory of finished ones: every result equals the alone result."""
    print("case batch staggered", flush=True)
    alone, _ = run_mix(batch.BATCHED, 2 * GB)
    big = max(batch.footprint(batch.FDTD, 40, 31, 17, 3), batch.footprint(batch.GROUPBY, 200_000, 65535))
    with batch.Executor(int(1.5 * big) + 3 * batch.footprint(batch.GMM, 30_000, 16, 32, 6), batch.BATCHED) as ex:
        ids = []
        for j, (kind, args) in enumerate(MIX):
            ids.append(submit(ex, kind, args, 100 + j))
            ex.tick()
            ex.tick()
        ex.drain()
        assert [ex.checksums(i) for i in ids] == alone


@pytest.mark.parametrize("mode", MODES)
def test_single_and_multi_iteration_runs_agree(hip, mode):
    """Iterations launched back to back (no wait between them) give what one-at-a-time ticks give."""
    print(f"case batch multi-iteration mode={mode}", flush=True)
    fdtd_only = [(k, a) for k, a in MIX if k != batch.GROUPBY]
    sums = []
    for stepwise in (True, False):
        with batch.Executor(2 * GB, mode) as ex:
            ids = [submit(ex, kind, args, 300 + j) for j, (kind, args) in enumerate(fdtd_only)]
            if stepwise:
                while any(ex.status(i)[0] < batch.DONE for i in ids):
                    ex.tick()
            else:
                n, _, _ = ex.run(1 << 20)
                # the shortest instance needs 3 steps: 3 iterations, or 1 of three steps
                assert n == (1 if mode == batch.BATCHED else 3)
                ex.drain()
            sums.append([ex.checksums(i) for i in ids])
    assert sums[0] == sums[1]
