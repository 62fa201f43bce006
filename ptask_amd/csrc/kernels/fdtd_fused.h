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
