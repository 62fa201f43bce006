// fdtd_fused.h -- one full Yee step (H from old E, then E from new H) in a single pass.
//
// Traffic per cell-step: six fields read once, six written once (48 B), against 96 B for the
// six-task form.  Each cell's arithmetic is the expression of ptaskUnitTest/graphfdtd.cpp:185-312
// in the same order, so results are bit-identical to the six separate updates.
//
// Tiling.  A block owns a TI x TJ tile of (i,j) columns (thread (tx,ty) owns column
// (i0+tx, j0+ty)) and walks a slab of k planes upwards.  For every plane k it
//   1. computes new Hx,Hy,Hz of its own cells (old E at k and k+1; the plane k+1 values read
//      here are the plane k values of the next iteration and come back from cache) and writes
//      them out;
//   2. computes the one-cell halo on the low side that new E needs from other tiles: new Hx, Hz
//      on row j0-1 and new Hy, Hz on column i0-1 (bit-identical to what the owning tile writes);
//   3. after one barrier, computes new Ex,Ey,Ez from LDS (i-1 and j-1 neighbours) and from the
//      new Hx, Hy of plane k-1 that the thread carries.
// LDS is double-buffered by plane parity, so one barrier per plane suffices.
//
// The union domain is i in [0,Nx], j in [0,Ny], k in [0,Nz]; every load and store checks the
// extent of its own field (see include/ptk_fdtd.h).
#pragma once

#include <hip/hip_runtime.h>

#include "ptk_fdtd.h"

namespace ptk_fdtd {

struct Fields {
    const float *__restrict__ hx;
    const float *__restrict__ hy;
    const float *__restrict__ hz;
    const float *__restrict__ ex;
    const float *__restrict__ ey;
    const float *__restrict__ ez;
    float *__restrict__ ohx;
    float *__restrict__ ohy;
    float *__restrict__ ohz;
    float *__restrict__ oex;
    float *__restrict__ oey;
    float *__restrict__ oez;
};

// Flat offsets for each field's extents (i fastest), in 32-bit arithmetic: every field array
// must hold fewer than 2^31 elements (checked by the kernels before any access).
struct Ix {
    int Nx, Ny, Nz;
    __device__ int hx(int i, int j, int k) const { return i + (Nx + 1) * (j + Ny * k); }
    __device__ int hy(int i, int j, int k) const { return i + Nx * (j + (Ny + 1) * k); }
    __device__ int hz(int i, int j, int k) const { return i + Nx * (j + Ny * k); }
    __device__ int ex(int i, int j, int k) const { return i + Nx * (j + (Ny + 1) * k); }
    __device__ int ey(int i, int j, int k) const { return i + (Nx + 1) * (j + Ny * k); }
    __device__ int ez(int i, int j, int k) const { return i + (Nx + 1) * (j + (Ny + 1) * k); }
    // extents (graphfdtd.cpp:70-86)
    __device__ bool vhx(int i, int j, int k) const { return i <= Nx && j < Ny && k < Nz; }
    __device__ bool vhy(int i, int j, int k) const { return i < Nx && j <= Ny && k < Nz; }
    __device__ bool vhz(int i, int j, int k) const { return i < Nx && j < Ny && k <= Nz; }
    __device__ bool vex(int i, int j, int k) const { return i < Nx && j <= Ny && k <= Nz; }
    __device__ bool vey(int i, int j, int k) const { return i <= Nx && j < Ny && k <= Nz; }
    __device__ bool vez(int i, int j, int k) const { return i <= Nx && j <= Ny && k < Nz; }
};

struct Coef {
    float cH, cE, Cx, Cy, Cz;
};

__device__ inline Coef make_coef(const FDTD_PARAMS &p) {
    Coef c;
    c.cH = p.Dt / p.mu0;
    c.cE = p.Dt / p.eps0;
    c.Cx = p.Cx;
    c.Cy = p.Cy;
    c.Cz = p.Cz;
    return c;
}

// New Hx at (i,j,k) from old fields (graphfdtd.cpp:185-201).  Caller checks vhx.
__device__ inline float new_hx(const Fields &F, const Ix &X, const Coef &c, int i, int j, int k) {
    const float ey0 = F.ey[X.ey(i, j, k)], ey1 = F.ey[X.ey(i, j, k + 1)];
    const float ez0 = F.ez[X.ez(i, j, k)], ez1 = F.ez[X.ez(i, j + 1, k)];
    return F.hx[X.hx(i, j, k)] + c.cH * ((ey1 - ey0) * c.Cz - (ez1 - ez0) * c.Cy);
}
// New Hy (graphfdtd.cpp:204-221).  Caller checks vhy.
__device__ inline float new_hy(const Fields &F, const Ix &X, const Coef &c, int i, int j, int k) {
    const float ez0 = F.ez[X.ez(i, j, k)], ez1 = F.ez[X.ez(i + 1, j, k)];
    const float ex0 = F.ex[X.ex(i, j, k)], ex1 = F.ex[X.ex(i, j, k + 1)];
    return F.hy[X.hy(i, j, k)] + c.cH * ((ez1 - ez0) * c.Cx - (ex1 - ex0) * c.Cz);
}
// New Hz (graphfdtd.cpp:223-240).  Caller checks vhz.
__device__ inline float new_hz(const Fields &F, const Ix &X, const Coef &c, int i, int j, int k) {
    const float ex0 = F.ex[X.ex(i, j, k)], ex1 = F.ex[X.ex(i, j + 1, k)];
    const float ey0 = F.ey[X.ey(i, j, k)], ey1 = F.ey[X.ey(i + 1, j, k)];
    return F.hz[X.hz(i, j, k)] + c.cH * ((ex1 - ex0) * c.Cy - (ey1 - ey0) * c.Cx);
}

// TRUE when the largest field (the (Nx+1)(Ny+1)(Nz+1) union bounds them all) is addressable with
// 32-bit offsets.
__device__ inline bool fits_int32(const FDTD_PARAMS &p) {
    return (long long)(p.Nx + 1) * (p.Ny + 1) * (p.Nz + 1) < 2147483647ll;
}

// One tile of one time step over the k slab [k0, k1) of the union domain.
// LDS layout: [parity][field][TJ+1][TI+1], index [ty+1][tx+1] for own cells, row 0 and column 0
// for the low-side halo.
template <int TI, int TJ>
__device__ inline void fused_tile(const Fields &F, const FDTD_PARAMS &p, int i0, int j0, int k0, int k1,
                                  float (*lds)[3][TJ + 1][TI + 1]) {
    const Ix X{p.Nx, p.Ny, p.Nz};
    const Coef c = make_coef(p);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i = i0 + tx, j = j0 + ty;
    const bool own = i <= p.Nx && j <= p.Ny;  // inside the union domain
    const bool haloRow = ty == 0 && j0 >= 1 && i <= p.Nx;
    const bool haloCol = tx == 0 && i0 >= 1 && j <= p.Ny;

    // new Hx, Hy of this column on plane k0-1 (E at k0 needs them)
    float hxm = 0.f, hym = 0.f;
    if (own && k0 >= 1) {
        if (X.vhx(i, j, k0 - 1)) hxm = new_hx(F, X, c, i, j, k0 - 1);
        if (X.vhy(i, j, k0 - 1)) hym = new_hy(F, X, c, i, j, k0 - 1);
    }
    for (int k = k0; k < k1; ++k) {
        float (*s)[TJ + 1][TI + 1] = lds[k & 1];
        float hx = 0.f, hy = 0.f, hz = 0.f;
        if (own) {
            if (X.vhx(i, j, k)) {
                hx = new_hx(F, X, c, i, j, k);
                F.ohx[X.hx(i, j, k)] = hx;
            }
            if (X.vhy(i, j, k)) {
                hy = new_hy(F, X, c, i, j, k);
                F.ohy[X.hy(i, j, k)] = hy;
            }
            if (X.vhz(i, j, k)) {
                hz = new_hz(F, X, c, i, j, k);
                F.ohz[X.hz(i, j, k)] = hz;
            }
            s[0][ty + 1][tx + 1] = hx;
            s[1][ty + 1][tx + 1] = hy;
            s[2][ty + 1][tx + 1] = hz;
        }
        if (haloRow) {  // row j0-1: Hx and Hz
            const int jm = j0 - 1;
            s[0][0][tx + 1] = X.vhx(i, jm, k) ? new_hx(F, X, c, i, jm, k) : 0.f;
            s[2][0][tx + 1] = X.vhz(i, jm, k) ? new_hz(F, X, c, i, jm, k) : 0.f;
        }
        if (haloCol) {  // column i0-1: Hy and Hz
            const int im = i0 - 1;
            s[1][ty + 1][0] = X.vhy(im, j, k) ? new_hy(F, X, c, im, j, k) : 0.f;
            s[2][ty + 1][0] = X.vhz(im, j, k) ? new_hz(F, X, c, im, j, k) : 0.f;
        }
        __syncthreads();
        if (own) {
            // Ex (graphfdtd.cpp:243-264)
            if (X.vex(i, j, k)) {
                const float e = F.ex[X.ex(i, j, k)];
                float r = e;
                if (!(j < 1 || k < 1 || j > p.Ny - 1 || k > p.Nz - 1))
                    r = e + c.cE * ((hz - s[2][ty][tx + 1]) * c.Cy - (hy - hym) * c.Cz);
                F.oex[X.ex(i, j, k)] = r;
            }
            // Ey (graphfdtd.cpp:267-288)
            if (X.vey(i, j, k)) {
                const float e = F.ey[X.ey(i, j, k)];
                float r = e;
                if (!(i < 1 || k < 1 || i > p.Nx - 1 || k > p.Nz - 1))
                    r = e + c.cE * ((hx - hxm) * c.Cz - (hz - s[2][ty + 1][tx]) * c.Cx);
                F.oey[X.ey(i, j, k)] = r;
            }
            // Ez (graphfdtd.cpp:291-312)
            if (X.vez(i, j, k)) {
                const float e = F.ez[X.ez(i, j, k)];
                float r = e;
                if (!(i < 1 || j < 1 || i > p.Nx - 1 || j > p.Ny - 1))
                    r = e + c.cE * ((hy - s[1][ty + 1][tx]) * c.Cx - (hx - s[0][ty][tx + 1]) * c.Cy);
                F.oez[X.ez(i, j, k)] = r;
            }
        }
        hxm = hx;
        hym = hy;
    }
}

}  // namespace ptk_fdtd

namespace ptk_fdtd {

// ---- overlapped tiles: one whole step, one barrier per plane --------------------------------
//
// A block of 64 x R threads (R = blockDim.y, 3..MAXR) covers columns i0-1 .. i0+62 (lane l ->
// i = i0-1+l) and rows j0-1 .. j0+R-2 (wave r -> j = j0-1+r) and owns the inner 62 x (R-2)
// cells: it writes new H and E there only.  Every thread computes new H of its cell, so the
// i-1 / j-1 neighbours that an owned cell's new E needs are in the block; on the low edge these
// are cells of the neighbouring tile, recomputed bit for bit.  The old E that new H needs at
// i+1 / j+1 is likewise in the block.  New H of a cell outside a field's extent, or on the
// block's high edge, is formed from zeros or unrelated values and never reaches an output: an
// owned cell's new E reads new H only at in-extent cells of the block (graphfdtd.cpp:70-86).
//
// Per plane k a thread loads its own old H(k) and old E(k+1) once -- the plane k+1 values are
// carried to the next plane in registers, and both are prefetched two planes ahead -- so HBM
// sees each input once per tile (48 B per cell-step plus the overlap).  Neighbour values pass
// through LDS, double-buffered by plane parity:
//   H phase   new H(k) from registers and E(k) neighbours in e[k&1];
//             write new H(k) to h[k&1] and old E(k+1) to e[(k+1)&1];
//   barrier   (LDS only: loads in flight stay in flight);
//   E phase   new E(k) from registers (old E(k), new H(k), new H(k-1)) and h[k&1] neighbours.
// e[(k+1)&1] was last read in the H phase of plane k-1 and h[k&1] in the E phase of plane k-2,
// both before the barrier of plane k-1, so one barrier per plane orders every access.
// A slab [k0, k1) starts one plane early (k0-1) to form new H(k0-1) for E(k0).
#ifndef PTK_OVERLAP_AHEAD
#define PTK_OVERLAP_AHEAD 2  // planes of loads in flight ahead of the plane being computed (1 or 2)
#endif
#ifndef PTK_OVERLAP_STORE_POLICY
#define PTK_OVERLAP_STORE_POLICY 0  // cache-policy bits of the output stores
#endif

// LDS planes with one padding row and column on each side, so that the four neighbours of
// (r, l) are at fixed offsets from its slot: index [r + 1][l + 1].
template <int MAXR>
struct OverlapLds {
    float e[2][3][MAXR + 2][66];  // old Ex, Ey, Ez of a plane
    float h[2][3][MAXR + 2][66];  // new Hx, Hy, Hz of a plane
};

// Barrier that orders LDS accesses only; outstanding global loads are not waited for.
__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// x-neighbours across the lanes of one wave (a tile row): lane l reads lane l + 1 through the
// data-parallel-primitive mux (wave_shl:1), a VALU operand modifier instead of an LDS read.
// Lane 63 reads 0; it is a halo lane of every tile.
__device__ inline float lane_next(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xF, 0xF, false));
}

// Raw buffer access to one field array.  A byte offset at or beyond the array's size reads 0
// and drops a store (the buffer unit's range check), so a lane outside a field's extent issues
// the same instruction with bit 31 of its offset set instead of branching around it: every
// plane then issues a fixed sequence of memory instructions and the loads of later planes stay
// in flight across the barrier.  Arrays must hold fewer than 2^31 bytes.
constexpr unsigned kNoCell = 0x80000000u;

struct FieldBuf {
    __amdgpu_buffer_rsrc_t rs;
    __device__ float load(unsigned off) const {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, (int)off, 0, 0));
    }
    __device__ void store(unsigned off, float v) const {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rs, (int)off, 0,
                                              PTK_OVERLAP_STORE_POLICY);
    }
};

__device__ inline FieldBuf field_buf(const float *p, long long elems) {
    return FieldBuf{__builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)(elems * 4), 0x00020000)};
}

// TRUE when every field array is under 2^31 bytes (the buffer range check of overlap_tile).
__device__ inline bool fits_buffer(const FDTD_PARAMS &p) {
    return (long long)(p.Nx + 1) * (p.Ny + 1) * (p.Nz + 1) < (1ll << 29);
}

template <int MAXR>
__device__ inline void overlap_tile(const Fields &F, const FDTD_PARAMS &p, int i0, int j0, int k0, int k1,
                                    OverlapLds<MAXR> &L) {
    const Ix X{p.Nx, p.Ny, p.Nz};
    const Coef c = make_coef(p);
    const int R = (int)blockDim.y;
    const int l = (int)threadIdx.x, r = (int)threadIdx.y;
    const int i = i0 - 1 + l, j = j0 - 1 + r;
    const int Nx = p.Nx, Ny = p.Ny, Nz = p.Nz;
    const bool dom = i >= 0 && i <= Nx && j >= 0 && j <= Ny;
    const bool own = dom && l >= 1 && l <= 62 && r >= 1 && r <= R - 2;
    // byte offset of (i, j, 0) in each field, bit 31 set when the column is outside the field
    auto col = [](bool ok, int idx) { return ok ? 4u * (unsigned)idx : kNoCell; };
    const unsigned qHx = col(dom && j < Ny, X.hx(i, j, 0)), qHy = col(dom && i < Nx, X.hy(i, j, 0));
    const unsigned qHz = col(dom && i < Nx && j < Ny, X.hz(i, j, 0)), qEx = col(dom && i < Nx, X.ex(i, j, 0));
    const unsigned qEy = col(dom && j < Ny, X.ey(i, j, 0)), qEz = col(dom, X.ez(i, j, 0));
    const unsigned ownBit = own ? 0u : kNoCell;
    // byte plane strides
    const unsigned sHx = 4u * (Nx + 1) * Ny, sHy = 4u * Nx * (Ny + 1), sHz = 4u * Nx * Ny;
    const unsigned sEx = 4u * Nx * (Ny + 1), sEy = 4u * (Nx + 1) * Ny, sEz = 4u * (Nx + 1) * (Ny + 1);
    const FieldBuf bHx = field_buf(F.hx, (long long)(Nx + 1) * Ny * Nz), bHy = field_buf(F.hy, (long long)Nx * (Ny + 1) * Nz),
                   bHz = field_buf(F.hz, (long long)Nx * Ny * (Nz + 1)), bEx = field_buf(F.ex, (long long)Nx * (Ny + 1) * (Nz + 1)),
                   bEy = field_buf(F.ey, (long long)(Nx + 1) * Ny * (Nz + 1)), bEz = field_buf(F.ez, (long long)(Nx + 1) * (Ny + 1) * Nz);
    const FieldBuf wHx = field_buf(F.ohx, (long long)(Nx + 1) * Ny * Nz), wHy = field_buf(F.ohy, (long long)Nx * (Ny + 1) * Nz),
                   wHz = field_buf(F.ohz, (long long)Nx * Ny * (Nz + 1)), wEx = field_buf(F.oex, (long long)Nx * (Ny + 1) * (Nz + 1)),
                   wEy = field_buf(F.oey, (long long)(Nx + 1) * Ny * (Nz + 1)), wEz = field_buf(F.oez, (long long)(Nx + 1) * (Ny + 1) * Nz);
    // offset of plane k of a column: the plane test is uniform over the block (a scalar mask)
    auto at = [](unsigned q, unsigned s, int k, bool ok) { return (q + (unsigned)k * s) | (ok ? 0u : kNoCell); };
    // per-lane parts of the face tests of graphfdtd.cpp:243-312
    const bool jIn = j >= 1 && j <= Ny - 1, iIn = i >= 1 && i <= Nx - 1;

    // Register sets S_k = {old E(k+1), old H(k)}.  Plane k reads S_{k-1} (its E part is old
    // E(k)) and S_k, while the loads of the next PTK_OVERLAP_AHEAD sets are in flight: the sets
    // rotate through an unrolled loop body, so no set is ever copied.  A set S_k with k >= k1 is
    // never used and its loads are dropped.
    struct Set {
        float ex, ey, ez, hx, hy, hz;
    };
    auto load_set = [&](Set &S, int k) {
        const bool use = k < k1;
        S.ex = bEx.load(at(qEx, sEx, k + 1, use && k + 1 <= Nz));
        S.ey = bEy.load(at(qEy, sEy, k + 1, use && k + 1 <= Nz));
        S.ez = bEz.load(at(qEz, sEz, k + 1, use && k + 1 < Nz));
        S.hx = bHx.load(at(qHx, sHx, k, use && k < Nz));
        S.hy = bHy.load(at(qHy, sHy, k, use && k < Nz));
        S.hz = bHz.load(at(qHz, sHz, k, use && k <= Nz));
    };
    const int ks = k0 >= 1 ? k0 - 1 : 0;
    Set S0, S1, S2, S3;
    S0.ex = bEx.load(at(qEx, sEx, ks, ks <= Nz));  // old E(ks)
    S0.ey = bEy.load(at(qEy, sEy, ks, ks <= Nz));
    S0.ez = bEz.load(at(qEz, sEz, ks, ks < Nz));
    S0.hx = S0.hy = S0.hz = 0.f;
    load_set(S1, ks);
#if PTK_OVERLAP_AHEAD == 2
    load_set(S2, ks + 1);
#endif
    const int y = r + 1, x = l + 1;  // this thread's LDS slot
    L.e[ks & 1][0][y][x] = S0.ex;
    L.e[ks & 1][1][y][x] = S0.ey;
    L.e[ks & 1][2][y][x] = S0.ez;
    float hxm = 0.f, hym = 0.f;  // new Hx, Hy of plane k-1
    lds_barrier();
    // plane k: P = S_{k-1}, S = S_k; loads S_{k+PTK_OVERLAP_AHEAD} into N
    auto plane = [&](int k, const Set &P, const Set &S, Set &N) {
        const int b = k & 1;
        load_set(N, k + PTK_OVERLAP_AHEAD);
        const float ex = P.ex, ey = P.ey, ez = P.ez;  // old E(k)
        const float ezj = L.e[b][2][y + 1][x], ezi = L.e[b][2][y][x + 1];
        const float exj = L.e[b][0][y + 1][x], eyi = L.e[b][1][y][x + 1];
        // graphfdtd.cpp:185-240, in the reference's operation order
        const float hx = S.hx + c.cH * ((S.ey - ey) * c.Cz - (ezj - ez) * c.Cy);
        const float hy = S.hy + c.cH * ((ezi - ez) * c.Cx - (S.ex - ex) * c.Cz);
        const float hz = S.hz + c.cH * ((exj - ex) * c.Cy - (eyi - ey) * c.Cx);
        L.h[b][0][y][x] = hx;
        L.h[b][1][y][x] = hy;
        L.h[b][2][y][x] = hz;
        L.e[b ^ 1][0][y][x] = S.ex;
        L.e[b ^ 1][1][y][x] = S.ey;
        L.e[b ^ 1][2][y][x] = S.ez;
        lds_barrier();
        const float hzj = L.h[b][2][y - 1][x], hzi = L.h[b][2][y][x - 1];
        const float hyi = L.h[b][1][y][x - 1], hxj = L.h[b][0][y - 1][x];
        // graphfdtd.cpp:243-312: interior cells update, tangential faces copy through
        const bool kIn = k >= 1 && k <= Nz - 1;
        const float wx = ex + c.cE * ((hz - hzj) * c.Cy - (hy - hym) * c.Cz);
        const float wy = ey + c.cE * ((hx - hxm) * c.Cz - (hz - hzi) * c.Cx);
        const float wz = ez + c.cE * ((hy - hyi) * c.Cx - (hx - hxj) * c.Cy);
        const float vx = (jIn && kIn) ? wx : ex;
        const float vy = (iIn && kIn) ? wy : ey;
        const float vz = (iIn && jIn) ? wz : ez;
        const bool out = k >= k0;
        wHx.store(at(qHx | ownBit, sHx, k, out && k < Nz), hx);
        wHy.store(at(qHy | ownBit, sHy, k, out && k < Nz), hy);
        wHz.store(at(qHz | ownBit, sHz, k, out && k <= Nz), hz);
        wEx.store(at(qEx | ownBit, sEx, k, out && k <= Nz), vx);
        wEy.store(at(qEy | ownBit, sEy, k, out && k <= Nz), vy);
        wEz.store(at(qEz | ownBit, sEz, k, out && k < Nz), vz);
        hxm = hx;
        hym = hy;
    };
#if PTK_OVERLAP_AHEAD == 2
    for (int k = ks; k < k1; k += 4) {
        plane(k, S0, S1, S3);
        if (k + 1 < k1) plane(k + 1, S1, S2, S0);
        if (k + 2 < k1) plane(k + 2, S2, S3, S1);
        if (k + 3 < k1) plane(k + 3, S3, S0, S2);
    }
#else
    (void)S3;
    for (int k = ks; k < k1; k += 3) {
        plane(k, S0, S1, S2);
        if (k + 1 < k1) plane(k + 1, S1, S2, S0);
        if (k + 2 < k1) plane(k + 2, S2, S0, S1);
    }
#endif
}

// ---- several time steps per pass on overlapped tiles -------------------------------------------
//
// A block of 64 x R threads (R = blockDim.y, 2T+1..MAXR) advances T time steps in one pass.  It
// covers columns i0-T .. i0+63-T (lane l -> i = i0-T+l) and rows j0-T .. j0+R-1-T (wave r ->
// j = j0-T+r) and owns the inner (64-2T) x (R-2T) cells, for which it writes H and E after the T
// steps.  Each half step loses one cell of valid data on one side (new H needs E at i+1 / j+1,
// new E needs H at i-1 / j-1), so after T steps the block minus T cells on each side is exact;
// the lost cells are recomputed by the neighbouring tiles, bit for bit.  HBM sees each input read
// and each output written once per T steps (48/T B per cell-step plus the overlap).
//
// Planes stream upwards; step s (1..T) runs 2(s-1) planes behind the first, on plane
// p_s = m - 2(s-1) in iteration m:
//   A  for each s: H_s(p_s) from E_{s-1}(p_s), E_{s-1}(p_s+1), H_{s-1}(p_s) and the E_{s-1}(p_s)
//      plane; write H_s(p_s) and E_0(m+1) to LDS
//   barrier
//   B  for each s: E_s(p_s) from E_{s-1}(p_s), H_s(p_s), H_s(p_s-1) and the H_s(p_s) plane; write
//      E_s(p_s) to LDS (s < T); store H_T, E_T of plane p_T if owned
//   barrier
// Step 0 is the input: E_0(m), E_0(m+1) and H_0(m) are loaded two planes ahead.  Step s-1's
// values two and one planes behind its own (those of planes p_s and p_s+1) are carried in
// registers (for T >= 3, the E values are read back from the own slots of the LDS planes).  Ez
// and Ey at i + 1 come from the next lane, so no Ey plane of E_0 is stored, nor, for T < 3, of
// any step (an Ey slot is only read by its own lane).  E_s
// planes are double-buffered by iteration parity (E_s(p_s) is written in B of iteration m and
// read as a plane in A of iteration m+2); H_s planes are written in A and read in B of the same
// iteration (HB = 1, two barriers per plane), or by parity (HB = 2, one barrier per plane).  A
// slab [k0, k1) starts T planes early and runs 2(T-1) planes late for the lag.  Every expression
// is the reference's, in its order (graphfdtd.cpp:185-312), so the result is bitwise that of T
// sequential steps.
template <int T, int MAXR, int HB>
struct OverlapNLds {
    float e[T][2][3][MAXR + 2][66];  // E_0 .. E_{T-1} planes by iteration parity
    float h[HB][T][3][MAXR + 2][66]; // H_1 .. H_T planes of the iteration (by parity if HB = 2)
};

// One field array for the multi-step tiles: the column's byte offset travels in a VGPR and the
// plane's in an SGPR, so a plane costs no vector address arithmetic.  A lane outside the field's
// columns carries bit 31 in its column offset, so its loads read 0 and its stores are dropped.
// Plane offsets are always those of planes inside the array, so a valid cell's column plus plane
// offset is below the array's size and no sum exceeds 2^32: the range check drops exactly the
// invalid lanes whether or not it counts the plane offset.
struct PlaneBuf {
    __amdgpu_buffer_rsrc_t rs;
    __device__ float load(unsigned col, unsigned plane) const {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, (int)col, (int)plane, 0));
    }
    __device__ void store(unsigned col, unsigned plane, float v) const {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rs, (int)col, (int)plane,
                                              PTK_OVERLAP_STORE_POLICY);
    }
};

__device__ inline PlaneBuf plane_buf(const float *p, unsigned bytes) {
    return PlaneBuf{__builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, 0x00020000)};
}

struct F3 {
    float x, y, z;
};

template <int T, int MAXR, int HB>
__device__ inline void overlapn_tile(const Fields &F, const FDTD_PARAMS &p, int i0, int j0, int k0, int k1,
                                     OverlapNLds<T, MAXR, HB> &L) {
    static_assert(T >= 1 && 2 * T < 64, "steps per pass");
    const Ix X{p.Nx, p.Ny, p.Nz};
    const Coef c = make_coef(p);
    const int R = (int)blockDim.y;
    const int l = (int)threadIdx.x, r = (int)threadIdx.y;
    const int i = i0 - T + l, j = j0 - T + r;
    const int Nx = p.Nx, Ny = p.Ny, Nz = p.Nz;
    const bool dom = i >= 0 && i <= Nx && j >= 0 && j <= Ny;
    const bool own = dom && l >= T && l <= 63 - T && r >= T && r <= R - 1 - T;
    auto col = [](bool ok, int idx) { return ok ? 4u * (unsigned)idx : kNoCell; };
    const unsigned qHx = col(dom && j < Ny, X.hx(i, j, 0)), qHy = col(dom && i < Nx, X.hy(i, j, 0));
    const unsigned qHz = col(dom && i < Nx && j < Ny, X.hz(i, j, 0)), qEx = col(dom && i < Nx, X.ex(i, j, 0));
    const unsigned qEy = col(dom && j < Ny, X.ey(i, j, 0)), qEz = col(dom, X.ez(i, j, 0));
    const unsigned ownBit = own ? 0u : kNoCell;  // stores also carry bit 31 where not owned
    const unsigned sHx = 4u * (Nx + 1) * Ny, sHy = 4u * Nx * (Ny + 1), sHz = 4u * Nx * Ny;
    const unsigned sEx = 4u * Nx * (Ny + 1), sEy = 4u * (Nx + 1) * Ny, sEz = 4u * (Nx + 1) * (Ny + 1);
    const PlaneBuf rHx = plane_buf(F.hx, sHx * Nz), rHy = plane_buf(F.hy, sHy * Nz), rHz = plane_buf(F.hz, sHz * (Nz + 1)),
                   rEx = plane_buf(F.ex, sEx * (Nz + 1)), rEy = plane_buf(F.ey, sEy * (Nz + 1)), rEz = plane_buf(F.ez, sEz * Nz);
    const PlaneBuf wHx = plane_buf(F.ohx, sHx * Nz), wHy = plane_buf(F.ohy, sHy * Nz), wHz = plane_buf(F.ohz, sHz * (Nz + 1)),
                   wEx = plane_buf(F.oex, sEx * (Nz + 1)), wEy = plane_buf(F.oey, sEy * (Nz + 1)), wEz = plane_buf(F.oez, sEz * Nz);
    const bool jIn = j >= 1 && j <= Ny - 1, iIn = i >= 1 && i <= Nx - 1;

    struct Set {
        F3 e, h;
    };
    // S_m = {E_0(m+1), H_0(m)}.  The owned planes below k1 depend on input planes up to T above
    // them (each step reaches one plane up), so only m <= k1 + T - 2 feed owned cells; a later
    // set reloads that set's planes (cached) and a plane past a field's last is read as its
    // last: those values reach no stored cell, as the planes formed from them do not.
    auto load_set = [&](Set &S, int m) {
        const int k = m < k1 + T - 2 ? m : k1 + T - 2;
        const int ke = k + 1 <= Nz ? k + 1 : Nz, kez = k + 1 <= Nz - 1 ? k + 1 : Nz - 1;
        const int kh = k <= Nz - 1 ? k : Nz - 1, khz = k <= Nz ? k : Nz;
        S.e.x = rEx.load(qEx, (unsigned)ke * sEx);
        S.e.y = rEy.load(qEy, (unsigned)ke * sEy);
        S.e.z = rEz.load(qEz, (unsigned)kez * sEz);
        S.h.x = rHx.load(qHx, (unsigned)kh * sHx);
        S.h.y = rHy.load(qHy, (unsigned)kh * sHy);
        S.h.z = rHz.load(qHz, (unsigned)khz * sHz);
    };
    // six dropped stores: the prologue then leaves the memory counter as a loop iteration does,
    // so the waits the compiler places in the loop stay two planes behind the loads
    auto pad_stores = [&]() {
        wHx.store(kNoCell, 0u, 0.f);
        wHy.store(kNoCell, 0u, 0.f);
        wHz.store(kNoCell, 0u, 0.f);
        wEx.store(kNoCell, 0u, 0.f);
        wEy.store(kNoCell, 0u, 0.f);
        wEz.store(kNoCell, 0u, 0.f);
    };
    constexpr int kLag = 2 * (T - 1);
    constexpr bool kEFromLds = T >= 3;  // registers: 128 per lane at 16 waves per CU
    // Owned planes from k0 need step s from plane k0 - (T - s) (E) or one below (H), i.e. the
    // first step from k0 - T; the last owned plane k1 - 1 leaves step T at iteration k1 - 1 +
    // 2(T-1).  Iterations m in [ms, me).
    const int ms = k0 >= T ? k0 - T : 0, me = k1 + kLag;
    Set S0, S1, S2, S3;
    {
        const int ke = ms, kez = ms <= Nz - 1 ? ms : Nz - 1;  // ms <= k0 < k1 <= Nz + 1
        S0.e.x = rEx.load(qEx, (unsigned)ke * sEx);  // E_0(ms)
        S0.e.y = rEy.load(qEy, (unsigned)ke * sEy);
        S0.e.z = rEz.load(qEz, (unsigned)kez * sEz);
        S0.h = F3{0.f, 0.f, 0.f};
    }
    load_set(S1, ms);
    pad_stores();
    load_set(S2, ms + 1);
    pad_stores();
    const int y = r + 1, x = l + 1;
    L.e[0][ms & 1][0][y][x] = S0.e.x;
    L.e[0][ms & 1][2][y][x] = S0.e.z;
    // carried own values of steps 1..T: planes one and two behind the step's own (index 0 unused)
    F3 h1[T + 1], h2[T + 1], e1[T + 1], e2[T + 1];
#pragma unroll
    for (int s = 0; s <= T; ++s) h1[s] = h2[s] = e1[s] = e2[s] = F3{0.f, 0.f, 0.f};

    lds_barrier();
    auto plane = [&](int m, const Set &P, const Set &S, Set &N) {
        const int b = m & 1, hb = HB == 2 ? b : 0;
        load_set(N, m + 2);
        // issue the loads of the plane two ahead before this plane's LDS traffic and its waits
        __builtin_amdgcn_sched_barrier(0);
        // ---- A: H_s(p_s) for every step
        F3 hn[T + 1], eo[T + 1];
#pragma unroll
        for (int s = 1; s <= T; ++s) {
            // E_{s-1}(p_s) and E_{s-1}(p_s + 1): with three or more steps, the own slots of the
            // E_{s-1} planes in LDS rather than registers carried across planes
            const auto &E = L.e[s - 1][b];
            const auto &Eu = L.e[s - 1][b ^ 1];
            const F3 ep = s == 1 ? P.e : kEFromLds ? F3{E[0][y][x], E[1][y][x], E[2][y][x]} : e2[s - 1];
            const F3 eu = s == 1 ? S.e : kEFromLds ? F3{Eu[0][y][x], Eu[1][y][x], Eu[2][y][x]} : e1[s - 1];
            const F3 ho = s == 1 ? S.h : h2[s - 1];  // H_{s-1}(p_s)
            eo[s] = ep;
            hn[s].x = ho.x + c.cH * ((eu.y - ep.y) * c.Cz - (E[2][y + 1][x] - ep.z) * c.Cy);
            // Ez and Ey at i + 1 from the next lane (the same values its LDS slots hold)
            hn[s].y = ho.y + c.cH * ((lane_next(ep.z) - ep.z) * c.Cx - (eu.x - ep.x) * c.Cz);
            hn[s].z = ho.z + c.cH * ((E[0][y + 1][x] - ep.x) * c.Cy - (lane_next(ep.y) - ep.y) * c.Cx);
            L.h[hb][s - 1][0][y][x] = hn[s].x;
            L.h[hb][s - 1][1][y][x] = hn[s].y;
            L.h[hb][s - 1][2][y][x] = hn[s].z;
        }
        L.e[0][b ^ 1][0][y][x] = S.e.x;
        L.e[0][b ^ 1][2][y][x] = S.e.z;
        lds_barrier();
        // ---- B: E_s(p_s) for every step (graphfdtd.cpp:243-312: interior cells update, faces
        // copy through)
        F3 en[T + 1];
#pragma unroll
        for (int s = 1; s <= T; ++s) {
            const int ps = m - 2 * (s - 1);
            const bool kIn = ps >= 1 && ps <= Nz - 1;
            const F3 e = eo[s];    // E_{s-1}(p_s)
            const F3 hp = h1[s];  // H_s(p_s - 1)
            const auto &H = L.h[hb][s - 1];
            const float wx = e.x + c.cE * ((hn[s].z - H[2][y - 1][x]) * c.Cy - (hn[s].y - hp.y) * c.Cz);
            const float wy = e.y + c.cE * ((hn[s].x - hp.x) * c.Cz - (hn[s].z - H[2][y][x - 1]) * c.Cx);
            const float wz = e.z + c.cE * ((hn[s].y - H[1][y][x - 1]) * c.Cx - (hn[s].x - H[0][y - 1][x]) * c.Cy);
            en[s].x = (jIn && kIn) ? wx : e.x;
            en[s].y = (iIn && kIn) ? wy : e.y;
            en[s].z = (iIn && jIn) ? wz : e.z;
            if (s < T) {  // E_s(p_s) in this parity's buffer was last read in A
                L.e[s][b][0][y][x] = en[s].x;
                if (kEFromLds) L.e[s][b][1][y][x] = en[s].y;  // read only at its own slot
                L.e[s][b][2][y][x] = en[s].z;
            }
        }
        // plane q of each output if it is in the slab and the field, else bit 31 and plane 0
        const int q = m - kLag;
        const bool out = q >= k0 && q < k1;
        const bool o1 = out && q < Nz, o2 = out && q <= Nz;
        const unsigned m1 = ownBit | (o1 ? 0u : kNoCell), m2 = ownBit | (o2 ? 0u : kNoCell);
        const unsigned p1 = o1 ? (unsigned)q : 0u, p2 = o2 ? (unsigned)q : 0u;
        wHx.store(qHx | m1, p1 * sHx, hn[T].x);
        wHy.store(qHy | m1, p1 * sHy, hn[T].y);
        wHz.store(qHz | m2, p2 * sHz, hn[T].z);
        wEx.store(qEx | m2, p2 * sEx, en[T].x);
        wEy.store(qEy | m2, p2 * sEy, en[T].y);
        wEz.store(qEz | m1, p1 * sEz, en[T].z);
#pragma unroll
        for (int s = 1; s <= T; ++s) {
            h2[s] = h1[s];
            h1[s] = hn[s];
            e2[s] = e1[s];
            e1[s] = en[s];
        }
        // With one H buffer, A of the next plane must not overwrite it before every wave has
        // read it here.  With two, the next A writes the other buffer, and every plane B writes
        // here is read two planes on, after the next barrier; the E_0 plane A writes next (the
        // other parity of e[0]) was last read in this plane's A, before the barrier above.
        if (HB == 1) lds_barrier();
    };
    // The sets rotate through four planes, and the loop runs whole groups of four with no branch
    // inside: up to three planes past me - 1 are formed and discarded (q >= k1: nothing stored).
    for (int m = ms; m < me; m += 4) {
        plane(m, S0, S1, S3);
        plane(m + 1, S1, S2, S0);
        plane(m + 2, S2, S3, S1);
        plane(m + 3, S3, S0, S2);
    }
}

}  // namespace ptk_fdtd
