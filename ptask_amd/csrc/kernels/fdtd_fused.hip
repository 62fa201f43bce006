// fdtd_fused.hip -- one whole Yee time step per launch (code object fdtd_fused.hsaco).
//
// Task kernel ABI (a fused FDTD graph: one task per step instead of six):
//   FdtdStepFused(Hx, Hy, Hz, Ex, Ey, Ez, FDTD_PARAMS p, oHx, oHy, oHz, oEx, oEy, oEz)
// formal parameters 0-5 the current fields, 6 the constants, 7-12 the next fields.
// Launch: block exactly (64, TJ, 1); grid (ceil((Nx+1)/64), ceil((Ny+1)/TJ), slabs).  The grid
// may be smaller than the tile count: blocks stride over tiles and split [0, Nz] into
// gridDim.z slabs.  Compiled with -ffp-contract=off: bit-identical to the six updates.
#include "fdtd_common.h"
#include "fdtd_fused.h"

using namespace ptk_fdtd;

template <int TI, int TJ>
__device__ inline void fused_grid(const Fields &F, const FDTD_PARAMS &p) {
    __shared__ float lds[2][3][TJ + 1][TI + 1];
    if (blockDim.x != TI || blockDim.y != TJ || blockDim.z != 1) return;  // geometry contract
    if (!fits_int32(p)) return;  // the launcher rejects such grids before launching
    const int ti = (p.Nx + 1 + TI - 1) / TI, tj = (p.Ny + 1 + TJ - 1) / TJ;
    const int depth = p.Nz + 1;
    const int slabs = (int)gridDim.z;
    const int k0 = (int)((long long)blockIdx.z * depth / slabs), k1 = (int)((long long)(blockIdx.z + 1) * depth / slabs);
    for (int by = blockIdx.y; by < tj; by += gridDim.y)
        for (int bx = blockIdx.x; bx < ti; bx += gridDim.x) {
            fused_tile<TI, TJ>(F, p, bx * TI, by * TJ, k0, k1, lds);
            __syncthreads();  // the next tile reuses LDS
        }
}

#define FUSED_KERNEL(NAME, TI, TJ)                                                              \
    extern "C" __global__ void __launch_bounds__(TI * TJ)                                       \
    NAME(const float *hx, const float *hy, const float *hz, const float *ex, const float *ey,   \
         const float *ez, FDTD_PARAMS p, float *ohx, float *ohy, float *ohz, float *oex,         \
         float *oey, float *oez) {                                                               \
        const Fields F{hx, hy, hz, ex, ey, ez, ohx, ohy, ohz, oex, oey, oez};                   \
        fused_grid<TI, TJ>(F, p);                                                               \
    }

FUSED_KERNEL(FdtdStepFused4, 64, 4)
FUSED_KERNEL(FdtdStepFused, 64, 8)
FUSED_KERNEL(FdtdStepFused2, 64, 2)

// Overlapped-tile whole step (see overlap_tile in fdtd_fused.h), same ABI as FdtdStepFused.
// Launch: block (64, R, 1) with 3 <= R <= 16 (FdtdStepOverlap) or 3 <= R <= 8
// (FdtdStepOverlap8, half the LDS); grid (ceil((Nx+1)/62), ceil((Ny+1)/(R-2)), slabs).
// Smaller grids stride over the tiles; gridDim.z splits [0, Nz] into slabs.  LDS per block:
// 2 planes x 6 fields x (MAXR+2) x 66 floats (31.7 KB for the 8-row kernel).
#ifndef PTK_OVERLAP_XCD
#define PTK_OVERLAP_XCD 1  // XCD-contiguous tile order
#endif

template <int MAXR>
__device__ inline void overlap_grid(const Fields &F, const FDTD_PARAMS &p, OverlapLds<MAXR> &L) {
    const int R = (int)blockDim.y;
    if (blockDim.x != 64 || R < 3 || R > MAXR || blockDim.z != 1) return;  // geometry contract
    if (!fits_buffer(p)) return;  // the launcher rejects such grids before launching
    const int ti = (p.Nx + 1 + 61) / 62, tj = (p.Ny + 1 + R - 3) / (R - 2);
    // Tiles that share cache lines (i- and j-neighbours of one slab) go to the same XCD at about
    // the same time: workgroups are dealt to the 8 XCDs round-robin, so the linear block id is
    // renumbered to give each XCD a contiguous run of the (x fastest, then y, then z) tile order.
    const unsigned nblk = gridDim.x * gridDim.y * gridDim.z;
    unsigned b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    if (PTK_OVERLAP_XCD && nblk >= kNumXcd) {
        const unsigned xcd = b % kNumXcd, local = b / kNumXcd, per = nblk / kNumXcd, rem = nblk % kNumXcd;
        b = xcd * per + (xcd < rem ? xcd : rem) + local;
    }
    const int gx = (int)(b % gridDim.x), gy = (int)(b / gridDim.x % gridDim.y), gz = (int)(b / (gridDim.x * gridDim.y));
    const int depth = p.Nz + 1, slabs = (int)gridDim.z;
    const int k0 = (int)((long long)gz * depth / slabs), k1 = (int)((long long)(gz + 1) * depth / slabs);
    if (k0 >= k1) return;  // the whole block
    for (int by = gy; by < tj; by += gridDim.y)
        for (int bx = gx; bx < ti; bx += gridDim.x) {
            overlap_tile<MAXR>(F, p, bx * 62, by * (R - 2), k0, k1, L);
            lds_barrier();  // the next tile reuses LDS
        }
}

#define OVERLAP_KERNEL(NAME, MAXR)                                                                   \
    extern "C" __global__ void __launch_bounds__(64 * MAXR)                                         \
    NAME(const float *hx, const float *hy, const float *hz, const float *ex, const float *ey,         \
         const float *ez, FDTD_PARAMS p, float *ohx, float *ohy, float *ohz, float *oex, float *oey, \
         float *oez) {                                                                                \
        __shared__ OverlapLds<MAXR> L;                                                                \
        const Fields F{hx, hy, hz, ex, ey, ez, ohx, ohy, ohz, oex, oey, oez};                         \
        overlap_grid<MAXR>(F, p, L);                                                                  \
    }

OVERLAP_KERNEL(FdtdStepOverlap, 16)
OVERLAP_KERNEL(FdtdStepOverlap8, 8)

// T time steps per launch on overlapped tiles (overlapn_tile in fdtd_fused.h), same ABI as
// FdtdStepFused: the outputs hold the fields T steps after the inputs.  Launch: block (64, R, 1)
// with 2T+1 <= R <= 16; grid (ceil((Nx+1)/(64-2T)), ceil((Ny+1)/(R-2T)), slabs).  Smaller grids
// stride over the tiles; gridDim.z splits [0, Nz] into slabs.  Tiles are taken in the
// XCD-contiguous order of FdtdStepOverlap.
//   FdtdTwoStepOverlap    T = 2: 60 x (R-4) owned columns per block, 85.5 KB of LDS
//   FdtdThreeStepOverlap  T = 3: 58 x (R-6) owned columns per block, 125.3 KB of LDS
template <int T, int HB>
__device__ inline void overlapn_grid(const Fields &F, const FDTD_PARAMS &p, OverlapNLds<T, 16, HB> &L) {
    const int R = (int)blockDim.y;
    if (blockDim.x != 64 || R < 2 * T + 1 || R > 16 || blockDim.z != 1) return;  // geometry contract
    // the launcher rejects such grids before launching; planes are addressed from 0 to Nz
    if (!fits_buffer(p) || p.Nx < 1 || p.Ny < 1 || p.Nz < 1) return;
    const int ow = 64 - 2 * T, oh = R - 2 * T;  // owned columns and rows of a tile
    const int ti = (p.Nx + ow) / ow, tj = (p.Ny + oh) / oh;
    const unsigned nblk = gridDim.x * gridDim.y * gridDim.z;
    unsigned b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    if (PTK_OVERLAP_XCD && nblk >= kNumXcd) {
        const unsigned xcd = b % kNumXcd, local = b / kNumXcd, per = nblk / kNumXcd, rem = nblk % kNumXcd;
        b = xcd * per + (xcd < rem ? xcd : rem) + local;
    }
    const int gx = (int)(b % gridDim.x), gy = (int)(b / gridDim.x % gridDim.y), gz = (int)(b / (gridDim.x * gridDim.y));
    const int depth = p.Nz + 1, slabs = (int)gridDim.z;
    const int k0 = (int)((long long)gz * depth / slabs), k1 = (int)((long long)(gz + 1) * depth / slabs);
    if (k0 >= k1) return;  // the whole block
    for (int by = gy; by < tj; by += gridDim.y)
        for (int bx = gx; bx < ti; bx += gridDim.x) {
            overlapn_tile<T, 16>(F, p, bx * ow, by * oh, k0, k1, L);
            lds_barrier();  // the next tile reuses LDS
        }
}

#define OVERLAPN_KERNEL(NAME, T, HB)                                                                  \
    extern "C" __global__ void __launch_bounds__(1024)                                                \
    NAME(const float *hx, const float *hy, const float *hz, const float *ex, const float *ey,         \
         const float *ez, FDTD_PARAMS p, float *ohx, float *ohy, float *ohz, float *oex, float *oey, \
         float *oez) {                                                                                \
        __shared__ OverlapNLds<T, 16, HB> L;                                                          \
        const Fields F{hx, hy, hz, ex, ey, ez, ohx, ohy, ohz, oex, oey, oez};                         \
        overlapn_grid<T, HB>(F, p, L);                                                                \
    }

// One H buffer and two barriers per plane: measured faster than H planes by parity with one
// barrier (the 16 waves then issue each plane's loads together).
OVERLAPN_KERNEL(FdtdTwoStepOverlap, 2, 1)
OVERLAPN_KERNEL(FdtdThreeStepOverlap, 3, 1)

// The same T steps with a persistent 1-D grid of G blocks (one per CU): the N tiles split into
// q = N / G whole-depth tiles per block, taken by all blocks together from plane 0 upwards (the
// tiles live at the same time share their i/j overlap in L2), and the planes of the remaining
// N - qG tiles cut into G equal runs in (tile, plane) order, one per block.  Every block then
// streams q(Nz+1) + (N - qG)(Nz+1)/G planes plus the pipeline's lag once or twice, instead of
// whole slabs in rounds of G blocks.  Any G covers every tile and plane exactly once.  Launch:
// block (64, R, 1), grid (G, 1, 1).
template <int T, int HB>
__device__ inline void overlapn_persistent(const Fields &F, const FDTD_PARAMS &p, OverlapNLds<T, 16, HB> &L) {
    const int R = (int)blockDim.y;
    if (blockDim.x != 64 || R < 2 * T + 1 || R > 16 || blockDim.z != 1) return;  // geometry contract
    if (gridDim.y != 1 || gridDim.z != 1) return;
    if (!fits_buffer(p) || p.Nx < 1 || p.Ny < 1 || p.Nz < 1) return;
    const int ow = 64 - 2 * T, oh = R - 2 * T;
    const int ti = (p.Nx + ow) / ow, tj = (p.Ny + oh) / oh, N = ti * tj, D = p.Nz + 1;
    const int G = (int)gridDim.x;
    int b = (int)blockIdx.x;
    if (PTK_OVERLAP_XCD && G >= kNumXcd) {  // XCD x takes blocks [x G/8, (x+1) G/8): neighbouring tiles
        const int xcd = b % kNumXcd, local = b / kNumXcd, per = G / kNumXcd, rem = G % kNumXcd;
        b = xcd * per + (xcd < rem ? xcd : rem) + local;
    }
    const int q = N / G;
    const long long W = (long long)(N - q * G) * D;  // tile-planes left after the whole tiles
    long long w = W * b / G;
    const long long hi = W * (b + 1) / G;
    // one call site for both kinds of segment (the tile body is large)
    for (int r = 0;; ++r) {
        int t, k0, k1;
        if (r < q) {  // whole-depth tile
            t = r * G + b, k0 = 0, k1 = D;
        } else if (w < hi) {  // a run of the remaining tiles' planes, within one tile
            t = q * G + (int)(w / D), k0 = (int)(w % D);
            k1 = (int)(k0 + (hi - w) < D ? k0 + (hi - w) : D);
            w += k1 - k0;
        } else {
            break;
        }
        overlapn_tile<T, 16>(F, p, (t % ti) * ow, (t / ti) * oh, k0, k1, L);
        lds_barrier();  // the next segment reuses LDS
    }
}

extern "C" __global__ void __launch_bounds__(1024)
FdtdThreeStepPersistent(const float *hx, const float *hy, const float *hz, const float *ex, const float *ey,
                        const float *ez, FDTD_PARAMS p, float *ohx, float *ohy, float *ohz, float *oex,
                        float *oey, float *oez) {
    __shared__ OverlapNLds<3, 16, 1> L;
    const Fields F{hx, hy, hz, ex, ey, ez, ohx, ohy, ohz, oex, oey, oez};
    overlapn_persistent<3, 1>(F, p, L);
}

// ---- half-step kernels: all three H updates in one streaming pass, then all three E updates ----
//
// FdtdHFields(Hx, Hy, Hz, Ex, Ey, Ez, FDTD_PARAMS p, oHx, oHy, oHz)
// FdtdEFields(Ex, Ey, Ez, Hx, Hy, Hz, FDTD_PARAMS p, oEx, oEy, oEz)   (Hx,Hy,Hz already new)
// Per cell-step: 6 reads + 3 writes each, 72 B for the pair.  The column-walk decomposition of
// fdtd_common.h over the union domain's (Nx+1)*(Ny+1) columns: any launch geometry covers every
// cell once; the k-neighbour of the walked direction is carried in registers, so there is no
// barrier and no LDS.
#include "fdtd_common.h"
#include "fdtd_halves.h"

extern "C" __global__ void __launch_bounds__(1024)
FdtdHFields(const float *__restrict__ Hx, const float *__restrict__ Hy, const float *__restrict__ Hz,
            const float *__restrict__ Ex, const float *__restrict__ Ey, const float *__restrict__ Ez,
            FDTD_PARAMS p, float *__restrict__ oHx, float *__restrict__ oHy, float *__restrict__ oHz) {
    if (!fits_int32(p)) return;
    const Sweep s = make_sweep(p.Nz + 1);
    if (s.k0 >= s.k1) return;
    const HalfH F{Hx, Hy, Hz, Ex, Ey, Ez, oHx, oHy, oHz};
    const Ix X{p.Nx, p.Ny, p.Nz};
    const Coef c = make_coef(p);
    const unsigned ncol = (unsigned)(p.Nx + 1) * (unsigned)(p.Ny + 1);
    for (unsigned q = s.lane; q < ncol; q += s.nlanes) h_column(F, X, c, q, s.k0, s.k1);
}

extern "C" __global__ void __launch_bounds__(1024)
FdtdEFields(const float *__restrict__ Ex, const float *__restrict__ Ey, const float *__restrict__ Ez,
            const float *__restrict__ Hx, const float *__restrict__ Hy, const float *__restrict__ Hz,
            FDTD_PARAMS p, float *__restrict__ oEx, float *__restrict__ oEy, float *__restrict__ oEz) {
    if (!fits_int32(p)) return;
    const Sweep s = make_sweep(p.Nz + 1);
    if (s.k0 >= s.k1) return;
    const HalfE F{Ex, Ey, Ez, Hx, Hy, Hz, oEx, oEy, oEz};
    const Ix X{p.Nx, p.Ny, p.Nz};
    const Coef c = make_coef(p);
    const unsigned ncol = (unsigned)(p.Nx + 1) * (unsigned)(p.Ny + 1);
    for (unsigned q = s.lane; q < ncol; q += s.nlanes) e_column(F, X, c, q, s.k0, s.k1);
}
