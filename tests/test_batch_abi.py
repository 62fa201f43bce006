"""The batched executor's C ABI on the CPU: the library loads, exports every function
include/ptk_batch.h declares, and computes instance footprints (no device calls)."""
import numpy as np

from ptask_amd import batch, fdtd, gmm, groupby


def test_library_exports_every_declared_symbol():
    L = batch.lib()
    names = batch.declared_symbols()
    assert len(names) == 23
    for n in names:
        assert hasattr(L, n), n


def test_footprints():
    up = lambda x: (x + 255) // 256 * 256
    nx, ny, nz = 17, 12, 9
    want = 2 * sum(up(4 * fdtd.elems(f, nx, ny, nz)) for f in range(6))
    assert batch.footprint(batch.FDTD, nx, ny, nz) == want
    rows, nk = 1000, 99
    K = nk + 1
    # one bucket, one chunk, at most two tiles; each array 256-aligned: head 256, btot, bstart,
    # tstart, tinfo, nz, berr, hist 256 each, kcnt 1024, c2 2048, s2 4096, tval 4096, tlow 1024
    assert groupby.scratch_bytes(rows, nk) == 14336
    assert batch.footprint(batch.GROUPBY, rows, nk) == (4 * up(4 * rows) + up(8 * (3 * K + 1)) + up(8 * K) +
                                                        up(8 * (3 + 3 * K)) + up(groupby.scratch_bytes(rows, nk)))
    n, d, k, steps = 5000, 7, 5, 9
    blocks = max(1, (-(-n // 128) + 15) // 16)
    model = gmm.model_elems(d, k, steps)
    assert batch.footprint(batch.GMM, n, d, k, steps) == (up(4 * n * d) + up(8 * model) +
                                                          up(8 * blocks * gmm.partial_stride(d, k)) + up(4 * k))
