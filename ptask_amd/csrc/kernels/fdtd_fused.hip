// fdtd_fused.hip -- one whole Yee time step per launch (code object fdtd_fused.hsaco).
//
// Task kernel ABI (a fused FDTD graph: one task per step instead of six):
//   FdtdStepFused(Hx, Hy, Hz, Ex, Ey, Ez, FDTD_PARAMS p, oHx, oHy, oHz, oEx, oEy, oEz)
// formal parameters 0-5 the current fields, 6 the constants, 7-12 the next fields.
// Launch: block exactly (64, TJ, 1); grid (ceil((Nx+1)/64), ceil((Ny+1)/TJ), slabs).  The grid
// may be smaller than the tile count: blocks stride over tiles and split [0, Nz] into
// gridDim.z slabs.  Compiled with -ffp-contract=off: bit-identical to the six updates.
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

// ---- half-step kernels: all three H updates in one streaming pass, then all three E updates ----
//
// FdtdHFields(Hx, Hy, Hz, Ex, Ey, Ez, FDTD_PARAMS p, oHx, oHy, oHz)
// FdtdEFields(Ex, Ey, Ez, Hx, Hy, Hz, FDTD_PARAMS p, oEx, oEy, oEz)   (Hx,Hy,Hz already new)
// Per cell-step: 6 reads + 3 writes each, 72 B for the pair.  The column-walk decomposition of
// fdtd_common.h over the union domain's (Nx+1)*(Ny+1) columns: any launch geometry covers every
// cell once; the k-neighbour of the walked direction is carried in registers, so there is no
// barrier and no LDS.
#include "fdtd_common.h"

extern "C" __global__ void __launch_bounds__(1024)
FdtdHFields(const float *__restrict__ Hx, const float *__restrict__ Hy, const float *__restrict__ Hz,
            const float *__restrict__ Ex, const float *__restrict__ Ey, const float *__restrict__ Ez,
            FDTD_PARAMS p, float *__restrict__ oHx, float *__restrict__ oHy, float *__restrict__ oHz) {
    if (!fits_int32(p)) return;
    const int Nx = p.Nx, Ny = p.Ny, Nz = p.Nz;
    const Sweep s = make_sweep(Nz + 1);
    const Ix X{Nx, Ny, Nz};
    const Coef c = make_coef(p);
    const unsigned ncol = (unsigned)(Nx + 1) * (unsigned)(Ny + 1);
    for (unsigned q = s.lane; q < ncol; q += s.nlanes) {
        if (s.k0 >= s.k1) break;
        const int j = (int)(q / (unsigned)(Nx + 1)), i = (int)q - j * (Nx + 1);
        const bool hasEx = i < Nx, hasEy = j < Ny, hasEz = true;
        const bool hasEzI = i < Nx, hasEzJ = j < Ny;      // Ez(i+1,j), Ez(i,j+1) exist
        const bool hasExJ = i < Nx && j < Ny, hasEyI = i < Nx && j < Ny;
        float ex0 = hasEx ? Ex[X.ex(i, j, s.k0)] : 0.f;
        float ey0 = hasEy ? Ey[X.ey(i, j, s.k0)] : 0.f;
        for (int k = s.k0; k < s.k1; ++k) {
            const bool kin = k < Nz;  // planes with Hx, Hy and Ez
            const float ex1 = (hasEx && kin) ? Ex[X.ex(i, j, k + 1)] : 0.f;
            const float ey1 = (hasEy && kin) ? Ey[X.ey(i, j, k + 1)] : 0.f;
            const float ez0 = (hasEz && kin) ? Ez[X.ez(i, j, k)] : 0.f;
            if (X.vhx(i, j, k)) {
                const float ezj = Ez[X.ez(i, j + 1, k)];
                oHx[X.hx(i, j, k)] = Hx[X.hx(i, j, k)] + c.cH * ((ey1 - ey0) * c.Cz - (ezj - ez0) * c.Cy);
            }
            if (X.vhy(i, j, k)) {
                const float ezi = Ez[X.ez(i + 1, j, k)];
                oHy[X.hy(i, j, k)] = Hy[X.hy(i, j, k)] + c.cH * ((ezi - ez0) * c.Cx - (ex1 - ex0) * c.Cz);
            }
            if (X.vhz(i, j, k)) {
                const float exj = Ex[X.ex(i, j + 1, k)], eyi = Ey[X.ey(i + 1, j, k)];
                oHz[X.hz(i, j, k)] = Hz[X.hz(i, j, k)] + c.cH * ((exj - ex0) * c.Cy - (eyi - ey0) * c.Cx);
            }
            (void)hasEzI; (void)hasEzJ; (void)hasExJ; (void)hasEyI;
            ex0 = ex1;
            ey0 = ey1;
        }
    }
}

extern "C" __global__ void __launch_bounds__(1024)
FdtdEFields(const float *__restrict__ Ex, const float *__restrict__ Ey, const float *__restrict__ Ez,
            const float *__restrict__ Hx, const float *__restrict__ Hy, const float *__restrict__ Hz,
            FDTD_PARAMS p, float *__restrict__ oEx, float *__restrict__ oEy, float *__restrict__ oEz) {
    if (!fits_int32(p)) return;
    const int Nx = p.Nx, Ny = p.Ny, Nz = p.Nz;
    const Sweep s = make_sweep(Nz + 1);
    const Ix X{Nx, Ny, Nz};
    const Coef c = make_coef(p);
    const unsigned ncol = (unsigned)(Nx + 1) * (unsigned)(Ny + 1);
    for (unsigned q = s.lane; q < ncol; q += s.nlanes) {
        if (s.k0 >= s.k1) break;
        const int j = (int)(q / (unsigned)(Nx + 1)), i = (int)q - j * (Nx + 1);
        // new Hx(i,j,k-1), Hy(i,j,k-1) carried along the walk
        float hxm = (s.k0 >= 1 && X.vhx(i, j, s.k0 - 1)) ? Hx[X.hx(i, j, s.k0 - 1)] : 0.f;
        float hym = (s.k0 >= 1 && X.vhy(i, j, s.k0 - 1)) ? Hy[X.hy(i, j, s.k0 - 1)] : 0.f;
        for (int k = s.k0; k < s.k1; ++k) {
            const float hx = X.vhx(i, j, k) ? Hx[X.hx(i, j, k)] : 0.f;
            const float hy = X.vhy(i, j, k) ? Hy[X.hy(i, j, k)] : 0.f;
            if (X.vex(i, j, k)) {  // graphfdtd.cpp:243-264
                const float e = Ex[X.ex(i, j, k)];
                float r = e;
                if (!(j < 1 || k < 1 || j > Ny - 1 || k > Nz - 1))
                    r = e + c.cE * ((Hz[X.hz(i, j, k)] - Hz[X.hz(i, j - 1, k)]) * c.Cy - (hy - hym) * c.Cz);
                oEx[X.ex(i, j, k)] = r;
            }
            if (X.vey(i, j, k)) {  // graphfdtd.cpp:267-288
                const float e = Ey[X.ey(i, j, k)];
                float r = e;
                if (!(i < 1 || k < 1 || i > Nx - 1 || k > Nz - 1))
                    r = e + c.cE * ((hx - hxm) * c.Cz - (Hz[X.hz(i, j, k)] - Hz[X.hz(i - 1, j, k)]) * c.Cx);
                oEy[X.ey(i, j, k)] = r;
            }
            if (X.vez(i, j, k)) {  // graphfdtd.cpp:291-312
                const float e = Ez[X.ez(i, j, k)];
                float r = e;
                if (!(i < 1 || j < 1 || i > Nx - 1 || j > Ny - 1))
                    r = e + c.cE * ((hy - Hy[X.hy(i - 1, j, k)]) * c.Cx - (hx - Hx[X.hx(i, j - 1, k)]) * c.Cy);
                oEz[X.ez(i, j, k)] = r;
            }
            hxm = hx;
            hym = hy;
        }
    }
}
