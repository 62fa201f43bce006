/*
 * oracle/fdtd.c -- CPU restatement of the reference FDTD Yee updates.  TEST INFRASTRUCTURE ONLY:
 * this is the checker that HIP results are compared against; no product code links it.
 *
 * Follows ptaskUnitTest/graphfdtd.cpp operation for operation:
 *   fdtd_HxComputation_cpu :185-201   fdtd_HyComputation_cpu :204-221   fdtd_HzComputation_cpu :223-240
 *   fdtd_ExComputation_cpu :243-264   fdtd_EyComputation_cpu :267-288   fdtd_EzComputation_cpu :291-312
 *   fdtd_cpu step order    :314-340 (Hx, Hy, Hz from old E; then Ex, Ey, Ez from new H)
 * Index v(i,j,k) = i + j*d1 + k*d1*d2 (ptaskUnitTest/Simple3DArray.h:28-35).  The reference
 * nests its loops i, j, k (k innermost); cells of one update are independent, so the loops here
 * run k, j, i (unit stride innermost) with identical results and cache-friendly access.
 * Built with -ffp-contract=off: each cell is a fixed sequence of IEEE binary32 sub/mul/add, so
 * a device kernel doing the same sequence must agree bit for bit.
 *
 * The same code instantiated in binary64 is the "FP64 host run" whose distance to the binary32
 * result is reported beside the parity check.
 *
 * Pinned by (tests/test_oracle_fdtd.py): constant fields are a fixed point; linearity
 * step(a*F) = a*step(F) in binary64; faces copied through; brute-force per-cell re-derivation of
 * the curl from the Yee stencil on tiny grids through an independent numpy formulation; the
 * graph form (six host tasks) equals this direct form.
 */
#include <stdlib.h>
#include <stddef.h>
#include <string.h>

#include "ptk_oracle.h"

#define V(a, i, j, k, d1, d2) (a)[(size_t)(i) + (size_t)(j) * (size_t)(d1) + (size_t)(k) * (size_t)(d1) * (size_t)(d2)]

#define DEFINE_FDTD(T, SFX)                                                                        \
void ptk_oracle_fdtd_hx_##SFX(const T *Hx, const T *Ey, const T *Ez, const FDTD_PARAMS *p, T *o) { \
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;                                                  \
    const T c = (T)p->Dt / (T)p->mu0, Cy = (T)p->Cy, Cz = (T)p->Cz;                                \
    for (int k = 0; k < Nz; ++k)                                                               \
        for (int j = 0; j < Ny; ++j)                                                               \
            for (int i = 0; i < Nx + 1; ++i)                                                           \
                V(o, i, j, k, Nx + 1, Ny) = V(Hx, i, j, k, Nx + 1, Ny) +                           \
                    c * ((V(Ey, i, j, k + 1, Nx + 1, Ny) - V(Ey, i, j, k, Nx + 1, Ny)) * Cz -      \
                         (V(Ez, i, j + 1, k, Nx + 1, Ny + 1) - V(Ez, i, j, k, Nx + 1, Ny + 1)) * Cy); \
}                                                                                                  \
void ptk_oracle_fdtd_hy_##SFX(const T *Hy, const T *Ex, const T *Ez, const FDTD_PARAMS *p, T *o) { \
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;                                                  \
    const T c = (T)p->Dt / (T)p->mu0, Cx = (T)p->Cx, Cz = (T)p->Cz;                                \
    for (int k = 0; k < Nz; ++k)                                                                   \
        for (int j = 0; j < Ny + 1; ++j)                                                           \
            for (int i = 0; i < Nx; ++i)                                                           \
                V(o, i, j, k, Nx, Ny + 1) = V(Hy, i, j, k, Nx, Ny + 1) +                           \
                    c * ((V(Ez, i + 1, j, k, Nx + 1, Ny + 1) - V(Ez, i, j, k, Nx + 1, Ny + 1)) * Cx - \
                         (V(Ex, i, j, k + 1, Nx, Ny + 1) - V(Ex, i, j, k, Nx, Ny + 1)) * Cz);      \
}                                                                                                  \
void ptk_oracle_fdtd_hz_##SFX(const T *Hz, const T *Ex, const T *Ey, const FDTD_PARAMS *p, T *o) { \
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;                                                  \
    const T c = (T)p->Dt / (T)p->mu0, Cx = (T)p->Cx, Cy = (T)p->Cy;                                \
    for (int k = 0; k < Nz + 1; ++k)                                                                   \
        for (int j = 0; j < Ny; ++j)                                                               \
            for (int i = 0; i < Nx; ++i)                                                       \
                V(o, i, j, k, Nx, Ny) = V(Hz, i, j, k, Nx, Ny) +                                   \
                    c * ((V(Ex, i, j + 1, k, Nx, Ny + 1) - V(Ex, i, j, k, Nx, Ny + 1)) * Cy -      \
                         (V(Ey, i + 1, j, k, Nx + 1, Ny) - V(Ey, i, j, k, Nx + 1, Ny)) * Cx);      \
}                                                                                                  \
void ptk_oracle_fdtd_ex_##SFX(const T *Ex, const T *Hy, const T *Hz, const FDTD_PARAMS *p, T *o) { \
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;                                                  \
    const T c = (T)p->Dt / (T)p->eps0, Cy = (T)p->Cy, Cz = (T)p->Cz;                               \
    for (int k = 0; k < Nz + 1; ++k)                                                                   \
        for (int j = 0; j < Ny + 1; ++j)                                                           \
            for (int i = 0; i < Nx; ++i) {                                                     \
                if (j < 1 || k < 1 || j > Ny - 1 || k > Nz - 1) {                                  \
                    V(o, i, j, k, Nx, Ny + 1) = V(Ex, i, j, k, Nx, Ny + 1);                        \
                } else {                                                                           \
                    V(o, i, j, k, Nx, Ny + 1) = V(Ex, i, j, k, Nx, Ny + 1) +                       \
                        c * ((V(Hz, i, j, k, Nx, Ny) - V(Hz, i, j - 1, k, Nx, Ny)) * Cy -          \
                             (V(Hy, i, j, k, Nx, Ny + 1) - V(Hy, i, j, k - 1, Nx, Ny + 1)) * Cz);  \
                }                                                                                  \
            }                                                                                      \
}                                                                                                  \
void ptk_oracle_fdtd_ey_##SFX(const T *Ey, const T *Hx, const T *Hz, const FDTD_PARAMS *p, T *o) { \
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;                                                  \
    const T c = (T)p->Dt / (T)p->eps0, Cx = (T)p->Cx, Cz = (T)p->Cz;                               \
    for (int k = 0; k < Nz + 1; ++k)                                                               \
        for (int j = 0; j < Ny; ++j)                                                               \
            for (int i = 0; i < Nx + 1; ++i) {                                                     \
                if (i < 1 || k < 1 || i > Nx - 1 || k > Nz - 1) {                                  \
                    V(o, i, j, k, Nx + 1, Ny) = V(Ey, i, j, k, Nx + 1, Ny);                        \
                } else {                                                                           \
                    V(o, i, j, k, Nx + 1, Ny) = V(Ey, i, j, k, Nx + 1, Ny) +                       \
                        c * ((V(Hx, i, j, k, Nx + 1, Ny) - V(Hx, i, j, k - 1, Nx + 1, Ny)) * Cz -  \
                             (V(Hz, i, j, k, Nx, Ny) - V(Hz, i - 1, j, k, Nx, Ny)) * Cx);          \
                }                                                                                  \
            }                                                                                      \
}                                                                                                  \
void ptk_oracle_fdtd_ez_##SFX(const T *Ez, const T *Hx, const T *Hy, const FDTD_PARAMS *p, T *o) { \
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;                                                  \
    const T c = (T)p->Dt / (T)p->eps0, Cx = (T)p->Cx, Cy = (T)p->Cy;                               \
    for (int k = 0; k < Nz; ++k)                                                               \
        for (int j = 0; j < Ny + 1; ++j)                                                           \
            for (int i = 0; i < Nx + 1; ++i) {                                                         \
                if (i < 1 || j < 1 || i > Nx - 1 || j > Ny - 1) {                                  \
                    V(o, i, j, k, Nx + 1, Ny + 1) = V(Ez, i, j, k, Nx + 1, Ny + 1);                \
                } else {                                                                           \
                    V(o, i, j, k, Nx + 1, Ny + 1) = V(Ez, i, j, k, Nx + 1, Ny + 1) +               \
                        c * ((V(Hy, i, j, k, Nx, Ny + 1) - V(Hy, i - 1, j, k, Nx, Ny + 1)) * Cx -  \
                             (V(Hx, i, j, k, Nx + 1, Ny) - V(Hx, i, j - 1, k, Nx + 1, Ny)) * Cy);  \
                }                                                                                  \
            }                                                                                      \
}

DEFINE_FDTD(float, f32)
DEFINE_FDTD(double, f64)

/* One full step in place on six field arrays, using the scratch arrays for the out-of-place
 * results and copying back, exactly as fdtd_cpu swaps its array pointers (graphfdtd.cpp:314-340). */
#define DEFINE_STEPS(T, SFX)                                                                       \
void ptk_oracle_fdtd_run_##SFX(T *f[6], T *scratch[6], const FDTD_PARAMS *p, int steps) {          \
    for (int s = 0; s < steps; ++s) {                                                              \
        ptk_oracle_fdtd_hx_##SFX(f[0], f[4], f[5], p, scratch[0]);                                  \
        ptk_oracle_fdtd_hy_##SFX(f[1], f[3], f[5], p, scratch[1]);                                  \
        ptk_oracle_fdtd_hz_##SFX(f[2], f[3], f[4], p, scratch[2]);                                  \
        for (int h = 0; h < 3; ++h) { T *t = f[h]; f[h] = scratch[h]; scratch[h] = t; }             \
        ptk_oracle_fdtd_ex_##SFX(f[3], f[1], f[2], p, scratch[3]);                                  \
        ptk_oracle_fdtd_ey_##SFX(f[4], f[0], f[2], p, scratch[4]);                                  \
        ptk_oracle_fdtd_ez_##SFX(f[5], f[0], f[1], p, scratch[5]);                                  \
        for (int e = 3; e < 6; ++e) { T *t = f[e]; f[e] = scratch[e]; scratch[e] = t; }             \
    }                                                                                              \
}

DEFINE_STEPS(float, f32)
DEFINE_STEPS(double, f64)

/* Array-of-pointers forms for ctypes callers (the arrays are swapped, so the caller reads the
 * final fields back through f[] after the call). */
void ptk_oracle_fdtd_run_ptrs_f32(float **f, float **scratch, const FDTD_PARAMS *p, int steps) {
    ptk_oracle_fdtd_run_f32(f, scratch, p, steps);
}
void ptk_oracle_fdtd_run_ptrs_f64(double **f, double **scratch, const FDTD_PARAMS *p, int steps) {
    ptk_oracle_fdtd_run_f64(f, scratch, p, steps);
}

/* Fill field f of an Nx x Ny x Nz grid with the seeded U[-1,1) values the workload uses
 * (stream id = field id, element index = flat index). */
void ptk_oracle_fdtd_init_field(float *a, int field, int Nx, int Ny, int Nz, uint64_t seed) {
    const unsigned long long n = ptk_fdtd_field_elems(field, Nx, Ny, Nz);
    for (unsigned long long idx = 0; idx < n; ++idx)
        a[idx] = ptk_rand_sym_f32(seed, (uint64_t)field, idx);
}

/* ---- host-task ABI (ptask/hosttask.h:28-31): void f(UINT nArgs, void **ppArgs) -------------
 * Formal parameters in port order as graphfdtd wires them (graphfdtd.cpp:415-424):
 * args[0] own field, args[1..2] the two curl inputs, args[3] -> FDTD_PARAMS (a by-value struct
 * reaches a host task by pointer, ptask/hosttask.cpp:149-276), args[4] output. */
#define HOSTTASK(NAME, FN)                                                                         \
void NAME(unsigned int n, void **a) {                                                              \
    (void)n;                                                                                       \
    FN((const float *)a[0], (const float *)a[1], (const float *)a[2],                              \
       (const FDTD_PARAMS *)a[3], (float *)a[4]);                                                  \
}
HOSTTASK(HxComputation, ptk_oracle_fdtd_hx_f32)
HOSTTASK(HyComputation, ptk_oracle_fdtd_hy_f32)
HOSTTASK(HzComputation, ptk_oracle_fdtd_hz_f32)
HOSTTASK(ExComputation, ptk_oracle_fdtd_ex_f32)
HOSTTASK(EyComputation, ptk_oracle_fdtd_ey_f32)
HOSTTASK(EzComputation, ptk_oracle_fdtd_ez_f32)

/* Host-task forms of the fused graphs: the same updates grouped as the HIP task kernels group
 * them.  FdtdHFields(Hx,Hy,Hz,Ex,Ey,Ez,&p,oHx,oHy,oHz); FdtdEFields(Ex,Ey,Ez,Hx,Hy,Hz,&p,oEx,oEy,
 * oEz) with the new H; FdtdStepFused(Hx..Ez,&p,oHx..oEz) one whole step. */
void FdtdHFields(unsigned int n, void **a) {
    (void)n;
    const FDTD_PARAMS *p = (const FDTD_PARAMS *)a[6];
    ptk_oracle_fdtd_hx_f32((const float *)a[0], (const float *)a[4], (const float *)a[5], p, (float *)a[7]);
    ptk_oracle_fdtd_hy_f32((const float *)a[1], (const float *)a[3], (const float *)a[5], p, (float *)a[8]);
    ptk_oracle_fdtd_hz_f32((const float *)a[2], (const float *)a[3], (const float *)a[4], p, (float *)a[9]);
}

void FdtdEFields(unsigned int n, void **a) {
    (void)n;
    const FDTD_PARAMS *p = (const FDTD_PARAMS *)a[6];
    ptk_oracle_fdtd_ex_f32((const float *)a[0], (const float *)a[4], (const float *)a[5], p, (float *)a[7]);
    ptk_oracle_fdtd_ey_f32((const float *)a[1], (const float *)a[3], (const float *)a[5], p, (float *)a[8]);
    ptk_oracle_fdtd_ez_f32((const float *)a[2], (const float *)a[3], (const float *)a[4], p, (float *)a[9]);
}

void FdtdStepFused(unsigned int n, void **a) {
    (void)n;
    const FDTD_PARAMS *p = (const FDTD_PARAMS *)a[6];
    ptk_oracle_fdtd_hx_f32((const float *)a[0], (const float *)a[4], (const float *)a[5], p, (float *)a[7]);
    ptk_oracle_fdtd_hy_f32((const float *)a[1], (const float *)a[3], (const float *)a[5], p, (float *)a[8]);
    ptk_oracle_fdtd_hz_f32((const float *)a[2], (const float *)a[3], (const float *)a[4], p, (float *)a[9]);
    ptk_oracle_fdtd_ex_f32((const float *)a[3], (const float *)a[8], (const float *)a[9], p, (float *)a[10]);
    ptk_oracle_fdtd_ey_f32((const float *)a[4], (const float *)a[7], (const float *)a[9], p, (float *)a[11]);
    ptk_oracle_fdtd_ez_f32((const float *)a[5], (const float *)a[7], (const float *)a[8], p, (float *)a[12]);
}

/* The overlapped-tile whole-step kernels compute the same step (same arguments). */
void FdtdStepOverlap(unsigned int n, void **a) { FdtdStepFused(n, a); }
void FdtdStepOverlap8(unsigned int n, void **a) { FdtdStepFused(n, a); }

/* FdtdTwoStepOverlap / FdtdThreeStepOverlap / FdtdThreeStepPersistent(Hx..Ez,&p,oHx..oEz): T
 * whole steps, the
 * intermediate fields in temporary arrays. */
static void fdtd_steps(unsigned int n, void **a, int T) {
    const FDTD_PARAMS *p = (const FDTD_PARAMS *)a[6];
    const long long X = p->Nx, Y = p->Ny, Z = p->Nz;
    const long long sz[6] = {(X + 1) * Y * Z, X * (Y + 1) * Z, X * Y * (Z + 1),
                             X * (Y + 1) * (Z + 1), (X + 1) * Y * (Z + 1), (X + 1) * (Y + 1) * Z};
    float *buf[2][6];
    for (int f = 0; f < 6; ++f)
        for (int k = 0; k < 2; ++k) buf[k][f] = (float *)malloc((size_t)(sz[f] > 0 ? sz[f] : 1) * sizeof(float));
    for (int t = 0; t < T; ++t) {
        void *s[13];
        s[6] = a[6];
        for (int f = 0; f < 6; ++f) {
            s[f] = t == 0 ? a[f] : (void *)buf[(t - 1) & 1][f];
            s[7 + f] = t == T - 1 ? a[7 + f] : (void *)buf[t & 1][f];
        }
        FdtdStepFused(n, s);
    }
    for (int f = 0; f < 6; ++f)
        for (int k = 0; k < 2; ++k) free(buf[k][f]);
}

void FdtdTwoStepOverlap(unsigned int n, void **a) { fdtd_steps(n, a, 2); }
void FdtdThreeStepOverlap(unsigned int n, void **a) { fdtd_steps(n, a, 3); }
void FdtdThreeStepPersistent(unsigned int n, void **a) { fdtd_steps(n, a, 3); }
