// fdtd.hip -- the six Yee half-step updates as PTask HIP task kernels (code object fdtd.hsaco).
//
// Kernel ABI (what a graph built like ptaskUnitTest/graphfdtd.cpp:415-424 binds): formal
// parameters in port-index order, buffers as device pointers, the FDTD_PARAMS block by value:
//     XxComputation(float* own, float* in1, float* in2, FDTD_PARAMS p, float* out)
// Updates are out of place and follow the CPU restatement operation for operation
// (ptaskUnitTest/graphfdtd.cpp:185-312); this file must be compiled with -ffp-contract=off and
// correctly rounded division so that each cell is bit-identical to the host result.
//
// H updates cover the whole H array (graphfdtd.cpp:185-240).  E updates copy the tangential
// faces through unchanged (graphfdtd.cpp:243-312).  Memory traffic per cell: 3 reads + 1 write
// of 4 bytes; the k-direction neighbour is carried in a register along the column walk and the
// i/j neighbours are re-read from L1/L2 (see fdtd_common.h for the launch decomposition).

#include "fdtd_common.h"
#include "ptk_prng.h"

using ptk_fdtd::Sweep;
using ptk_fdtd::make_sweep;

#define PTK_FDTD_KERNEL extern "C" __global__ void __launch_bounds__(1024)

// Hx = Hx + (Dt/mu0)*((Ey(i,j,k+1)-Ey(i,j,k))*Cz - (Ez(i,j+1,k)-Ez(i,j,k))*Cy)
PTK_FDTD_KERNEL HxComputation(const float *__restrict__ Hx, const float *__restrict__ Ey,
                              const float *__restrict__ Ez, FDTD_PARAMS p,
                              float *__restrict__ out) {
    const int d1 = p.Nx + 1, d2 = p.Ny, d3 = p.Nz;
    const Sweep s = make_sweep(d3);
    const float c = p.Dt / p.mu0;
    const unsigned ncol = (unsigned)d1 * (unsigned)d2;
    const long long P = (long long)d1 * d2;                  // Hx and Ey share (Nx+1, Ny)
    const long long PEz = (long long)(p.Nx + 1) * (p.Ny + 1);
    for (unsigned q = s.lane; q < ncol; q += s.nlanes) {
        if (s.k0 >= s.k1) break;
        const int j = (int)(q / (unsigned)d1), i = (int)q - j * d1;
        const float *h = Hx + q + s.k0 * P;
        const float *ey = Ey + q + s.k0 * P;
        const float *ez = Ez + i + (long long)j * (p.Nx + 1) + s.k0 * PEz;
        float *o = out + q + s.k0 * P;
        float ey0 = ey[0];
#pragma unroll PTK_FDTD_UNROLL
        for (int k = s.k0; k < s.k1; ++k) {
            const float ey1 = ey[P];
            const float ez0 = ez[0], ez1 = ez[p.Nx + 1];
            o[0] = h[0] + c * ((ey1 - ey0) * p.Cz - (ez1 - ez0) * p.Cy);
            ey0 = ey1;
            h += P; ey += P; ez += PEz; o += P;
        }
    }
}

// Hy = Hy + (Dt/mu0)*((Ez(i+1,j,k)-Ez(i,j,k))*Cx - (Ex(i,j,k+1)-Ex(i,j,k))*Cz)
PTK_FDTD_KERNEL HyComputation(const float *__restrict__ Hy, const float *__restrict__ Ex,
                              const float *__restrict__ Ez, FDTD_PARAMS p,
                              float *__restrict__ out) {
    const int d1 = p.Nx, d2 = p.Ny + 1, d3 = p.Nz;
    const Sweep s = make_sweep(d3);
    const float c = p.Dt / p.mu0;
    const unsigned ncol = (unsigned)d1 * (unsigned)d2;
    const long long P = (long long)d1 * d2;                  // Hy and Ex share (Nx, Ny+1)
    const long long PEz = (long long)(p.Nx + 1) * (p.Ny + 1);
    for (unsigned q = s.lane; q < ncol; q += s.nlanes) {
        if (s.k0 >= s.k1) break;
        const int j = (int)(q / (unsigned)d1), i = (int)q - j * d1;
        const float *h = Hy + q + s.k0 * P;
        const float *ex = Ex + q + s.k0 * P;
        const float *ez = Ez + i + (long long)j * (p.Nx + 1) + s.k0 * PEz;
        float *o = out + q + s.k0 * P;
        float ex0 = ex[0];
#pragma unroll PTK_FDTD_UNROLL
        for (int k = s.k0; k < s.k1; ++k) {
            const float ex1 = ex[P];
            const float ez0 = ez[0], ez1 = ez[1];
            o[0] = h[0] + c * ((ez1 - ez0) * p.Cx - (ex1 - ex0) * p.Cz);
            ex0 = ex1;
            h += P; ex += P; ez += PEz; o += P;
        }
    }
}

// Hz = Hz + (Dt/mu0)*((Ex(i,j+1,k)-Ex(i,j,k))*Cy - (Ey(i+1,j,k)-Ey(i,j,k))*Cx)
PTK_FDTD_KERNEL HzComputation(const float *__restrict__ Hz, const float *__restrict__ Ex,
                              const float *__restrict__ Ey, FDTD_PARAMS p,
                              float *__restrict__ out) {
    const int d1 = p.Nx, d2 = p.Ny, d3 = p.Nz + 1;
    const Sweep s = make_sweep(d3);
    const float c = p.Dt / p.mu0;
    const unsigned ncol = (unsigned)d1 * (unsigned)d2;
    const long long P = (long long)d1 * d2;
    const long long PEx = (long long)p.Nx * (p.Ny + 1);
    const long long PEy = (long long)(p.Nx + 1) * p.Ny;
    for (unsigned q = s.lane; q < ncol; q += s.nlanes) {
        if (s.k0 >= s.k1) break;
        const int j = (int)(q / (unsigned)d1), i = (int)q - j * d1;
        const float *h = Hz + q + s.k0 * P;
        const float *ex = Ex + i + (long long)j * p.Nx + s.k0 * PEx;
        const float *ey = Ey + i + (long long)j * (p.Nx + 1) + s.k0 * PEy;
        float *o = out + q + s.k0 * P;
#pragma unroll PTK_FDTD_UNROLL
        for (int k = s.k0; k < s.k1; ++k) {
            const float ex0 = ex[0], ex1 = ex[p.Nx];
            const float ey0 = ey[0], ey1 = ey[1];
            o[0] = h[0] + c * ((ex1 - ex0) * p.Cy - (ey1 - ey0) * p.Cx);
            h += P; ex += PEx; ey += PEy; o += P;
        }
    }
}

// Ex interior (1<=j<=Ny-1, 1<=k<=Nz-1):
//   Ex = Ex + (Dt/eps0)*((Hz(i,j,k)-Hz(i,j-1,k))*Cy - (Hy(i,j,k)-Hy(i,j,k-1))*Cz); else copy.
PTK_FDTD_KERNEL ExComputation(const float *__restrict__ Ex, const float *__restrict__ Hy,
                              const float *__restrict__ Hz, FDTD_PARAMS p,
                              float *__restrict__ out) {
    const int d1 = p.Nx, d2 = p.Ny + 1, d3 = p.Nz + 1;
    const Sweep s = make_sweep(d3);
    const float c = p.Dt / p.eps0;
    const unsigned ncol = (unsigned)d1 * (unsigned)d2;
    const long long P = (long long)d1 * d2;                  // Ex and Hy share (Nx, Ny+1)
    const long long PHz = (long long)p.Nx * p.Ny;
    for (unsigned q = s.lane; q < ncol; q += s.nlanes) {
        if (s.k0 >= s.k1) break;
        const int j = (int)(q / (unsigned)d1), i = (int)q - j * d1;
        const float *e = Ex + q + s.k0 * P;
        float *o = out + q + s.k0 * P;
        if (j < 1 || j > p.Ny - 1) {
            for (int k = s.k0; k < s.k1; ++k, e += P, o += P) o[0] = e[0];
            continue;
        }
        // Hz(i,j,k) and Hz(i,j-1,k); Hy(i,j,k) and Hy(i,j,k-1) (carried).
        const float *hz = Hz + i + (long long)j * p.Nx + s.k0 * PHz;
        const float *hy = Hy + q + s.k0 * P;
        float hym = (s.k0 >= 1) ? hy[-P] : 0.0f;
#pragma unroll PTK_FDTD_UNROLL
        for (int k = s.k0; k < s.k1; ++k) {
            const float hy0 = (k <= p.Nz - 1) ? hy[0] : 0.0f;
            if (k < 1 || k > p.Nz - 1) {
                o[0] = e[0];
            } else {
                const float hz0 = hz[0], hzm = hz[-p.Nx];
                o[0] = e[0] + c * ((hz0 - hzm) * p.Cy - (hy0 - hym) * p.Cz);
            }
            hym = hy0;
            e += P; o += P; hz += PHz; hy += P;
        }
    }
}

// Ey interior (1<=i<=Nx-1, 1<=k<=Nz-1):
//   Ey = Ey + (Dt/eps0)*((Hx(i,j,k)-Hx(i,j,k-1))*Cz - (Hz(i,j,k)-Hz(i-1,j,k))*Cx); else copy.
PTK_FDTD_KERNEL EyComputation(const float *__restrict__ Ey, const float *__restrict__ Hx,
                              const float *__restrict__ Hz, FDTD_PARAMS p,
                              float *__restrict__ out) {
    const int d1 = p.Nx + 1, d2 = p.Ny, d3 = p.Nz + 1;
    const Sweep s = make_sweep(d3);
    const float c = p.Dt / p.eps0;
    const unsigned ncol = (unsigned)d1 * (unsigned)d2;
    const long long P = (long long)d1 * d2;                  // Ey and Hx share (Nx+1, Ny)
    const long long PHz = (long long)p.Nx * p.Ny;
    for (unsigned q = s.lane; q < ncol; q += s.nlanes) {
        if (s.k0 >= s.k1) break;
        const int j = (int)(q / (unsigned)d1), i = (int)q - j * d1;
        const float *e = Ey + q + s.k0 * P;
        float *o = out + q + s.k0 * P;
        if (i < 1 || i > p.Nx - 1) {
            for (int k = s.k0; k < s.k1; ++k, e += P, o += P) o[0] = e[0];
            continue;
        }
        const float *hx = Hx + q + s.k0 * P;
        const float *hz = Hz + i + (long long)j * p.Nx + s.k0 * PHz;
        float hxm = (s.k0 >= 1) ? hx[-P] : 0.0f;
#pragma unroll PTK_FDTD_UNROLL
        for (int k = s.k0; k < s.k1; ++k) {
            const float hx0 = (k <= p.Nz - 1) ? hx[0] : 0.0f;
            if (k < 1 || k > p.Nz - 1) {
                o[0] = e[0];
            } else {
                const float hz0 = hz[0], hzm = hz[-1];
                o[0] = e[0] + c * ((hx0 - hxm) * p.Cz - (hz0 - hzm) * p.Cx);
            }
            hxm = hx0;
            e += P; o += P; hx += P; hz += PHz;
        }
    }
}

// Ez interior (1<=i<=Nx-1, 1<=j<=Ny-1):
//   Ez = Ez + (Dt/eps0)*((Hy(i,j,k)-Hy(i-1,j,k))*Cx - (Hx(i,j,k)-Hx(i,j-1,k))*Cy); else copy.
PTK_FDTD_KERNEL EzComputation(const float *__restrict__ Ez, const float *__restrict__ Hx,
                              const float *__restrict__ Hy, FDTD_PARAMS p,
                              float *__restrict__ out) {
    const int d1 = p.Nx + 1, d2 = p.Ny + 1, d3 = p.Nz;
    const Sweep s = make_sweep(d3);
    const float c = p.Dt / p.eps0;
    const unsigned ncol = (unsigned)d1 * (unsigned)d2;
    const long long P = (long long)d1 * d2;
    const long long PHx = (long long)(p.Nx + 1) * p.Ny;
    const long long PHy = (long long)p.Nx * (p.Ny + 1);
    for (unsigned q = s.lane; q < ncol; q += s.nlanes) {
        if (s.k0 >= s.k1) break;
        const int j = (int)(q / (unsigned)d1), i = (int)q - j * d1;
        const float *e = Ez + q + s.k0 * P;
        float *o = out + q + s.k0 * P;
        if (i < 1 || i > p.Nx - 1 || j < 1 || j > p.Ny - 1) {
            for (int k = s.k0; k < s.k1; ++k, e += P, o += P) o[0] = e[0];
            continue;
        }
        const float *hy = Hy + i + (long long)j * p.Nx + s.k0 * PHy;
        const float *hx = Hx + i + (long long)j * (p.Nx + 1) + s.k0 * PHx;
#pragma unroll PTK_FDTD_UNROLL
        for (int k = s.k0; k < s.k1; ++k) {
            const float hy0 = hy[0], hym = hy[-1];
            const float hx0 = hx[0], hxm = hx[-(p.Nx + 1)];
            o[0] = e[0] + c * ((hy0 - hym) * p.Cx - (hx0 - hxm) * p.Cy);
            e += P; o += P; hy += PHy; hx += PHx;
        }
    }
}

// Seeded fill of one field array with the U[-1,1) values of include/ptk_prng.h: element n of
// field f is ptk_rand_sym_f32(seed, f, n), so a device-initialised grid equals a host-initialised
// one bit for bit.  Grid-stride over the `n` elements.
extern "C" __global__ void __launch_bounds__(256)
FdtdInitField(float *__restrict__ a, unsigned long long n, int field, unsigned long long seed) {
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long idx = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += stride)
        a[idx] = ptk_rand_sym_f32(seed, (uint64_t)field, idx);
}
