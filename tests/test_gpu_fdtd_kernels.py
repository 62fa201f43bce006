"""Parity of the six Yee task kernels (fdtd.hsaco) with the oracle, called through the C ABI.

Bar: bit-exact against the binary32 oracle (the kernels perform the same IEEE operations in the
same order); the <=1e-5 max relative error gate of the workload is asserted as well, and the
binary64 distance is reported.  Covers several launch geometries (the reference's
one-thread-per-cell CUDA geometry, the default column walk, odd shapes, a single block) and
degenerate grids with a dimension of 1.
"""
import numpy as np
import pytest

import oracle
from ptask_amd import fdtd as FD
from ptask_amd import paths

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mod(hip):
    return hip.Module(paths.kernel_path("fdtd.hsaco"))


def upload_fields(hip, F):
    bufs = []
    for a in F:
        b = hip.DeviceBuffer(a.nbytes)
        b.upload(a)
        bufs.append(b)
    return bufs


def run_steps(hip, mod, F, p, steps, geometry):
    nx, ny, nz = p.Nx, p.Ny, p.Nz
    cur = upload_fields(hip, F)
    nxt = [hip.DeviceBuffer(a.nbytes) for a in F]
    s = hip.Stream()
    for _ in range(steps):
        for half in ((0, 1, 2), (3, 4, 5)):
            for t in half:
                name, own, a, b = FD.TASKS[t]
                grid, block = geometry(*FD.extent(own, nx, ny, nz))
                args = FD.pack_task_args(hip, cur[own].ptr, cur[a].ptr, cur[b].ptr, p,
                                         nxt[own].ptr)
                hip.launch(mod.function(name), grid, block, args, s)
            for t in half:
                cur[t], nxt[t] = nxt[t], cur[t]
    s.synchronize()
    out = []
    for f, a in enumerate(F):
        h = np.empty_like(a)
        cur[f].download(h)
        out.append(h)
    s.destroy()
    return out


GEOMS = {
    "default": FD.default_geometry,
    "reference": FD.reference_geometry,
    "odd": lambda d1, d2, d3: ((5, 2, 3), (64, 3, 1)),
    "one_block": lambda d1, d2, d3: ((1, 1, 1), (64, 1, 1)),
    "thin_slabs": lambda d1, d2, d3: FD.default_geometry(d1, d2, d3, block=128, target_slab=1),
}


def assert_parity(got, ref, ref64=None):
    for f in range(6):
        assert got[f].shape == ref[f].shape
        if not np.array_equal(got[f].view(np.uint32), ref[f].view(np.uint32)):
            bad = np.flatnonzero(got[f] != ref[f])
            raise AssertionError(f"{FD.FIELDS[f]}: {bad.size} cells differ, first {bad[:5]}")
        rel = np.abs(got[f] - ref[f]).max() / max(np.abs(ref[f]).max(), 1e-30)
        assert rel <= 1e-5


@pytest.mark.parametrize("geom", sorted(GEOMS))
@pytest.mark.parametrize("shape", [(8, 8, 8), (13, 7, 5)])
def test_six_task_step_bitexact(hip, mod, shape, geom):
    nx, ny, nz = shape
    print(f"case shape={shape} geom={geom}", flush=True)
    p = oracle.make_params(nx, ny, nz, 0.5)
    F = oracle.init_fields(nx, ny, nz, seed=100 + nx)
    got = run_steps(hip, mod, F, FD.make_params(nx, ny, nz), 4, GEOMS[geom])
    ref = oracle.fdtd_run(F, p, 4)
    assert_parity(got, ref)


@pytest.mark.parametrize("shape", [(1, 1, 1), (1, 5, 3), (6, 1, 2), (4, 3, 1), (2, 2, 2)])
def test_degenerate_grids(hip, mod, shape):
    nx, ny, nz = shape
    print(f"case degenerate shape={shape}", flush=True)
    p = oracle.make_params(nx, ny, nz, 0.5)
    F = oracle.init_fields(nx, ny, nz, seed=7)
    got = run_steps(hip, mod, F, FD.make_params(nx, ny, nz), 3, FD.default_geometry)
    assert_parity(got, oracle.fdtd_run(F, p, 3))


def test_params_bytes_match_oracle():
    for shape in [(64, 64, 64), (384, 384, 384), (17, 33, 9)]:
        a = FD.make_params(*shape)
        b = oracle.make_params(*shape, 0.5)
        assert bytes(a) == bytes(b)


def test_long_run_64cubed_against_fp32_and_fp64(hip, mod):
    """64^3 for 10 steps: the CPU configuration of the workload, checked at the gate and
    reported against binary64."""
    nx = ny = nz = 64
    print("case 64^3 x 10 steps", flush=True)
    p = oracle.make_params(nx, ny, nz, 0.5)
    F = oracle.init_fields(nx, ny, nz, seed=2024)
    got = run_steps(hip, mod, F, FD.make_params(nx, ny, nz), 10, FD.default_geometry)
    ref = oracle.fdtd_run(F, p, 10)
    assert_parity(got, ref)
    ref64 = oracle.fdtd_run(F, p, 10, dtype=np.float64)
    for f in range(6):
        e64 = np.abs(got[f] - ref64[f]).max() / np.abs(ref64[f]).max()
        print(f"{FD.FIELDS[f]} max rel err vs binary64: {e64:.3e}")
        assert e64 < 1e-4


def test_device_init_matches_host_init(hip, mod):
    nx, ny, nz = 9, 6, 5
    F = oracle.init_fields(nx, ny, nz, seed=0xABCDEF)
    s = hip.Stream()
    for f, a in enumerate(F):
        b = hip.DeviceBuffer(a.nbytes)
        args = hip.ArgPack().ptr(b.ptr).u64(a.size).i32(f).u64(0xABCDEF)
        hip.launch(mod.function("FdtdInitField"), (3, 1, 1), (64, 1, 1), args, s)
        s.synchronize()
        h = np.empty_like(a)
        b.download(h)
        assert np.array_equal(h.view(np.uint32), a.view(np.uint32)), FD.FIELDS[f]


# ---- fused whole-step kernel (fdtd_fused.hsaco) ----------------------------------------------

FUSED = {"FdtdStepFused2": 2, "FdtdStepFused4": 4, "FdtdStepFused": 8}


def fused_geometry(nx, ny, nz, tj, slabs=None, max_tiles=None):
    ti = -(-(nx + 1) // 64)
    tjn = -(-(ny + 1) // tj)
    if max_tiles:
        ti, tjn = min(ti, max_tiles), min(tjn, max_tiles)
    if slabs is None:
        slabs = max(1, -(-(nz + 1) // 16))
    return (ti, tjn, slabs), (64, tj, 1)


def run_fused(hip, mod, name, F, p, steps, geo):
    cur = upload_fields(hip, F)
    nxt = [hip.DeviceBuffer(a.nbytes) for a in F]
    s = hip.Stream()
    for _ in range(steps):
        args = hip.ArgPack()
        for b in cur:
            args.ptr(b.ptr)
        args.raw(bytes(p), 4)
        for b in nxt:
            args.ptr(b.ptr)
        hip.launch(mod.function(name), geo[0], geo[1], args, s)
        cur, nxt = nxt, cur
    s.synchronize()
    out = []
    for f, a in enumerate(F):
        h = np.empty_like(a)
        cur[f].download(h)
        out.append(h)
    s.destroy()
    return out


@pytest.fixture(scope="module")
def fmod(hip):
    return hip.Module(paths.kernel_path("fdtd_fused.hsaco"))


@pytest.mark.parametrize("name", sorted(FUSED))
@pytest.mark.parametrize("shape", [(8, 8, 8), (70, 9, 5), (13, 21, 17), (1, 1, 1), (3, 1, 2)])
def test_fused_step_bitexact(hip, fmod, name, shape):
    nx, ny, nz = shape
    print(f"case fused {name} shape={shape}", flush=True)
    F = oracle.init_fields(nx, ny, nz, seed=31 + nz)
    p = FD.make_params(nx, ny, nz)
    got = run_fused(hip, fmod, name, F, p, 5, fused_geometry(nx, ny, nz, FUSED[name]))
    assert_parity(got, oracle.fdtd_run(F, oracle.make_params(nx, ny, nz, 0.5), 5))


@pytest.mark.parametrize("variant", ["strided_tiles", "one_plane_slabs", "single_slab"])
def test_fused_step_any_grid_covers_domain(hip, fmod, variant):
    nx, ny, nz = 150, 11, 9
    print(f"case fused grid variant={variant}", flush=True)
    geo = {"strided_tiles": fused_geometry(nx, ny, nz, 8, slabs=3, max_tiles=2),
           "one_plane_slabs": fused_geometry(nx, ny, nz, 8, slabs=nz + 1),
           "single_slab": fused_geometry(nx, ny, nz, 8, slabs=1)}[variant]
    F = oracle.init_fields(nx, ny, nz, seed=5)
    got = run_fused(hip, fmod, "FdtdStepFused", F, FD.make_params(nx, ny, nz), 3, geo)
    assert_parity(got, oracle.fdtd_run(F, oracle.make_params(nx, ny, nz, 0.5), 3))


def run_halves(hip, mod, F, p, steps, geometry):
    """FdtdHFields then FdtdEFields per step (72 B/cell-step form)."""
    nx, ny, nz = p.Nx, p.Ny, p.Nz
    cur = upload_fields(hip, F)
    nxt = [hip.DeviceBuffer(a.nbytes) for a in F]
    s = hip.Stream()
    grid, block = geometry(nx + 1, ny + 1, nz + 1)
    for _ in range(steps):
        a = hip.ArgPack()
        for f in (0, 1, 2, 3, 4, 5):
            a.ptr(cur[f].ptr)
        a.raw(bytes(p), 4)
        for f in (0, 1, 2):
            a.ptr(nxt[f].ptr)
        hip.launch(mod.function("FdtdHFields"), grid, block, a, s)
        e = hip.ArgPack()
        for f in (3, 4, 5):
            e.ptr(cur[f].ptr)
        for f in (0, 1, 2):
            e.ptr(nxt[f].ptr)
        e.raw(bytes(p), 4)
        for f in (3, 4, 5):
            e.ptr(nxt[f].ptr)
        hip.launch(mod.function("FdtdEFields"), grid, block, e, s)
        cur, nxt = nxt, cur
    s.synchronize()
    out = []
    for f, arr in enumerate(F):
        h = np.empty_like(arr)
        cur[f].download(h)
        out.append(h)
    s.destroy()
    return out


@pytest.mark.parametrize("geom", sorted(GEOMS))
@pytest.mark.parametrize("shape", [(8, 8, 8), (13, 7, 5), (1, 4, 1), (5, 1, 3)])
def test_half_step_kernels_bitexact(hip, fmod, shape, geom):
    nx, ny, nz = shape
    print(f"case halves shape={shape} geom={geom}", flush=True)
    F = oracle.init_fields(nx, ny, nz, seed=77 + ny)
    got = run_halves(hip, fmod, F, FD.make_params(nx, ny, nz), 4, GEOMS[geom])
    assert_parity(got, oracle.fdtd_run(F, oracle.make_params(nx, ny, nz, 0.5), 4))


def tile_geometry(tj, slab, max_tiles=None):
    """2-D blocks of 64 x tj over the columns with k slabs of ~slab planes; max_tiles caps the
    grid."""
    def geo(d1, d2, d3):
        ti, tjn = -(-d1 // 64), -(-d2 // tj)
        if max_tiles:
            ti, tjn = min(ti, max_tiles), min(tjn, max_tiles)
        return (ti, tjn, max(1, -(-d3 // slab))), (64, tj, 1)
    return geo


HALF_TILES = {
    "64x8_slab8": tile_geometry(8, 8),
    "64x4_slab1": tile_geometry(4, 1),
    "64x16_one_slab": tile_geometry(16, 1 << 20),
    "64x2_strided": tile_geometry(2, 3, max_tiles=2),
    "64x8_strided_single": tile_geometry(8, 5, max_tiles=1),
}


@pytest.mark.parametrize("geom", sorted(HALF_TILES))
@pytest.mark.parametrize("shape", [(70, 9, 11), (150, 20, 9), (63, 15, 2), (1, 1, 1)])
def test_half_step_2d_blocks_cover_domain(hip, fmod, shape, geom):
    """The column walk visits every (column, k) exactly once for 2-D blocks and capped grids."""
    nx, ny, nz = shape
    print(f"case halves tiles shape={shape} geom={geom}", flush=True)
    F = oracle.init_fields(nx, ny, nz, seed=11 + nx)
    got = run_halves(hip, fmod, F, FD.make_params(nx, ny, nz), 3, HALF_TILES[geom])
    assert_parity(got, oracle.fdtd_run(F, oracle.make_params(nx, ny, nz, 0.5), 3))


def overlap_geometry(rows, slab, max_tiles=None):
    """Overlapped tiles: blocks of 64 x rows own 62 x (rows-2) columns; k slabs of ~slab planes;
    max_tiles caps the grid so blocks take further tiles in grid strides."""
    def geo(nx, ny, nz):
        ti, tj = -(-(nx + 1) // 62), -(-(ny + 1) // (rows - 2))
        if max_tiles:
            ti, tj = min(ti, max_tiles), min(tj, max_tiles)
        return (ti, tj, max(1, -(-(nz + 1) // slab))), (64, rows, 1)
    return geo


OVERLAP = {
    "r16_slab48": (16, 48, None),
    "r8_slab4": (8, 4, None),
    "r3_slab1": (3, 1, None),
    "r4_one_slab": (4, 1 << 20, None),
    "r16_slab3_strided": (16, 3, 1),
    "r5_slab2_strided": (5, 2, 2),
}


@pytest.mark.parametrize("name", ["FdtdStepOverlap", "FdtdStepOverlap8"])
@pytest.mark.parametrize("geom", sorted(OVERLAP))
@pytest.mark.parametrize("shape", [(8, 8, 8), (61, 13, 4), (62, 14, 3), (124, 9, 7), (130, 29, 6),
                                   (13, 21, 17), (1, 1, 1), (3, 1, 2), (1, 5, 3)])
def test_overlap_step_bitexact(hip, fmod, shape, geom, name):
    """FdtdStepOverlap(8): whole steps on overlapped tiles equal the six reference updates bit for
    bit, for tile edges at and around the 62-column / (R-2)-row tile sizes and any slab split."""
    rows, slab, cap = OVERLAP[geom]
    if name == "FdtdStepOverlap8" and rows > 8:
        pytest.skip("the 8-row kernel takes blocks of at most 8 rows")
    nx, ny, nz = shape
    print(f"case overlap {name} shape={shape} geom={geom}", flush=True)
    F = oracle.init_fields(nx, ny, nz, seed=211 + nx + ny)
    geo = overlap_geometry(rows, slab, cap)(nx, ny, nz)
    got = run_fused(hip, fmod, name, F, FD.make_params(nx, ny, nz), 4, geo)
    assert_parity(got, oracle.fdtd_run(F, oracle.make_params(nx, ny, nz, 0.5), 4))


def overlapn_geometry(T, rows, slab, max_tiles=None):
    """T-step overlapped tiles: blocks of 64 x rows own (64-2T) x (rows-2T) columns; k slabs of
    ~slab planes; max_tiles caps the grid so blocks take further tiles in grid strides."""
    def geo(nx, ny, nz):
        ti, tj = -(-(nx + 1) // (64 - 2 * T)), -(-(ny + 1) // (rows - 2 * T))
        if max_tiles:
            ti, tj = min(ti, max_tiles), min(tj, max_tiles)
        return (ti, tj, max(1, -(-(nz + 1) // slab))), (64, rows, 1)
    return geo


MULTI_STEP = {"FdtdTwoStepOverlap": 2, "FdtdThreeStepOverlap": 3}
OVERLAPN = {
    "r16_slab16": (16, 16, None),
    "r8_slab4": (8, 4, None),
    "rmin_slab1": (0, 1, None),
    "rmin1_slab2": (1, 2, None),
    "r12_one_slab": (12, 1 << 20, None),
    "r16_slab3_strided": (16, 3, 1),
    "rmin2_slab5_strided": (2, 5, 2),
}


@pytest.mark.parametrize("name", sorted(MULTI_STEP))
@pytest.mark.parametrize("geom", sorted(OVERLAPN))
@pytest.mark.parametrize("shape", [(8, 8, 8), (57, 9, 4), (58, 10, 3), (59, 11, 4), (60, 12, 3),
                                   (119, 9, 7), (130, 29, 6), (13, 21, 17), (1, 1, 1), (3, 1, 2),
                                   (1, 5, 3), (2, 2, 9)])
def test_multi_step_overlap_bitexact(hip, fmod, shape, geom, name):
    """FdtdTwoStepOverlap / FdtdThreeStepOverlap: each launch advances T whole steps on
    overlapped tiles and equals T rounds of the six reference updates bit for bit, for tile edges
    at and around the (64-2T)-column / (R-2T)-row tile sizes, the smallest blocks (R = 2T+1),
    slabs of one plane up to the whole depth, and grids that stride over tiles."""
    T = MULTI_STEP[name]
    rows, slab, cap = OVERLAPN[geom]
    if rows < 2 * T + 1:  # "rmin*": the smallest block height for T, plus rows
        rows += 2 * T + 1
    nx, ny, nz = shape
    print(f"case {name} shape={shape} geom={geom}", flush=True)
    F = oracle.init_fields(nx, ny, nz, seed=307 + nx + ny)
    geo = overlapn_geometry(T, rows, slab, cap)(nx, ny, nz)
    got = run_fused(hip, fmod, name, F, FD.make_params(nx, ny, nz), 2, geo)
    assert_parity(got, oracle.fdtd_run(F, oracle.make_params(nx, ny, nz, 0.5), 2 * T))


@pytest.mark.parametrize("grid,rows", [(1, 16), (3, 7), (8, 16), (13, 11), (256, 16), (600, 16)])
@pytest.mark.parametrize("shape", [(8, 8, 8), (57, 9, 4), (130, 29, 6), (13, 21, 17), (1, 1, 1),
                                   (3, 1, 2), (200, 40, 3)])
def test_persistent_three_step_bitexact(hip, fmod, shape, grid, rows):
    """FdtdThreeStepPersistent: a 1-D grid of any size covers the whole-depth tiles and the even
    runs of the remaining tiles' planes exactly once; two launches equal six reference steps."""
    nx, ny, nz = shape
    print(f"case FdtdThreeStepPersistent shape={shape} grid={grid} rows={rows}", flush=True)
    F = oracle.init_fields(nx, ny, nz, seed=401 + nx + nz)
    got = run_fused(hip, fmod, "FdtdThreeStepPersistent", F, FD.make_params(nx, ny, nz), 2,
                    ((grid, 1, 1), (64, rows, 1)))
    assert_parity(got, oracle.fdtd_run(F, oracle.make_params(nx, ny, nz, 0.5), 6))
