// fdtd_halves.h -- the column walks of the two half-step passes, shared by the single-graph
// half-step kernels (fdtd_fused.hip) and the batched multi-instance kernels (batch.hip), so both
// produce bit-identical fields.
//
// h_column: new Hx, Hy, Hz of union-domain column q = i + j*(Nx+1) for k in [k0, k1), from the
//   old H and E (graphfdtd.cpp:185-240); the k+1 plane of Ex, Ey is carried in registers.
// e_column: new Ex, Ey, Ez of the same column from the old E and the NEW H
//   (graphfdtd.cpp:243-312, faces copied through); the k-1 plane of Hx, Hy is carried.
#pragma once

#include "fdtd_fused.h"

namespace ptk_fdtd {

struct HalfH {
    const float *__restrict__ hx;
    const float *__restrict__ hy;
    const float *__restrict__ hz;
    const float *__restrict__ ex;
    const float *__restrict__ ey;
    const float *__restrict__ ez;
    float *__restrict__ ohx;
    float *__restrict__ ohy;
    float *__restrict__ ohz;
};

struct HalfE {
    const float *__restrict__ ex;
    const float *__restrict__ ey;
    const float *__restrict__ ez;
    const float *__restrict__ hx;  // new H
    const float *__restrict__ hy;
    const float *__restrict__ hz;
    float *__restrict__ oex;
    float *__restrict__ oey;
    float *__restrict__ oez;
};

__device__ inline void h_column(const HalfH &F, const Ix &X, const Coef &c, unsigned q, int k0, int k1) {
    const int Nx = X.Nx, Ny = X.Ny, Nz = X.Nz;
    const int j = (int)(q / (unsigned)(Nx + 1)), i = (int)q - j * (Nx + 1);
    const bool hasEx = i < Nx, hasEy = j < Ny;
    float ex0 = hasEx ? F.ex[X.ex(i, j, k0)] : 0.f;
    float ey0 = hasEy ? F.ey[X.ey(i, j, k0)] : 0.f;
    for (int k = k0; k < k1; ++k) {
        const bool kin = k < Nz;  // planes with Hx, Hy and Ez
        const float ex1 = (hasEx && kin) ? F.ex[X.ex(i, j, k + 1)] : 0.f;
        const float ey1 = (hasEy && kin) ? F.ey[X.ey(i, j, k + 1)] : 0.f;
        const float ez0 = kin ? F.ez[X.ez(i, j, k)] : 0.f;
        if (X.vhx(i, j, k)) {
            const float ezj = F.ez[X.ez(i, j + 1, k)];
            F.ohx[X.hx(i, j, k)] = F.hx[X.hx(i, j, k)] + c.cH * ((ey1 - ey0) * c.Cz - (ezj - ez0) * c.Cy);
        }
        if (X.vhy(i, j, k)) {
            const float ezi = F.ez[X.ez(i + 1, j, k)];
            F.ohy[X.hy(i, j, k)] = F.hy[X.hy(i, j, k)] + c.cH * ((ezi - ez0) * c.Cx - (ex1 - ex0) * c.Cz);
        }
        if (X.vhz(i, j, k)) {
            const float exj = F.ex[X.ex(i, j + 1, k)], eyi = F.ey[X.ey(i + 1, j, k)];
            F.ohz[X.hz(i, j, k)] = F.hz[X.hz(i, j, k)] + c.cH * ((exj - ex0) * c.Cy - (eyi - ey0) * c.Cx);
        }
        ex0 = ex1;
        ey0 = ey1;
    }
}

__device__ inline void e_column(const HalfE &F, const Ix &X, const Coef &c, unsigned q, int k0, int k1) {
    const int Nx = X.Nx, Ny = X.Ny, Nz = X.Nz;
    const int j = (int)(q / (unsigned)(Nx + 1)), i = (int)q - j * (Nx + 1);
    // new Hx(i,j,k-1), Hy(i,j,k-1) carried along the walk
    float hxm = (k0 >= 1 && X.vhx(i, j, k0 - 1)) ? F.hx[X.hx(i, j, k0 - 1)] : 0.f;
    float hym = (k0 >= 1 && X.vhy(i, j, k0 - 1)) ? F.hy[X.hy(i, j, k0 - 1)] : 0.f;
    for (int k = k0; k < k1; ++k) {
        const float hx = X.vhx(i, j, k) ? F.hx[X.hx(i, j, k)] : 0.f;
        const float hy = X.vhy(i, j, k) ? F.hy[X.hy(i, j, k)] : 0.f;
        if (X.vex(i, j, k)) {  // graphfdtd.cpp:243-264
            const float e = F.ex[X.ex(i, j, k)];
            float r = e;
            if (!(j < 1 || k < 1 || j > Ny - 1 || k > Nz - 1))
                r = e + c.cE * ((F.hz[X.hz(i, j, k)] - F.hz[X.hz(i, j - 1, k)]) * c.Cy - (hy - hym) * c.Cz);
            F.oex[X.ex(i, j, k)] = r;
        }
        if (X.vey(i, j, k)) {  // graphfdtd.cpp:267-288
            const float e = F.ey[X.ey(i, j, k)];
            float r = e;
            if (!(i < 1 || k < 1 || i > Nx - 1 || k > Nz - 1))
                r = e + c.cE * ((hx - hxm) * c.Cz - (F.hz[X.hz(i, j, k)] - F.hz[X.hz(i - 1, j, k)]) * c.Cx);
            F.oey[X.ey(i, j, k)] = r;
        }
        if (X.vez(i, j, k)) {  // graphfdtd.cpp:291-312
            const float e = F.ez[X.ez(i, j, k)];
            float r = e;
            if (!(i < 1 || j < 1 || i > Nx - 1 || j > Ny - 1))
                r = e + c.cE * ((hy - F.hy[X.hy(i - 1, j, k)]) * c.Cx - (hx - F.hx[X.hx(i, j - 1, k)]) * c.Cy);
            F.oez[X.ez(i, j, k)] = r;
        }
        hxm = hx;
        hym = hy;
    }
}

}  // namespace ptk_fdtd
