// pca_blas.h -- the binary64 BLAS task kernels the reference's unrolled PCA graph binds
// (ptaskUnitTest/graphcupca.cpp:294-299: ptaskDcopy, ptaskDaxpy, ptaskDgemv, ptaskDnrm2,
// ptaskDscal, ptaskDger), with the by-value parameter structs of graphcupca.cpp:161-220.  The
// reference's CUDA kernels are not in its tree; each kernel here computes what the matching
// cpublasD* routine of ptaskUnitTest/pca_cpu.cpp:147-265 computes, in the same order of
// operations per result (one thread walks a dot product or a norm in index order), so a graph of
// these tasks equals pca_cpu bit for bit.
//
// Argument order is the ports' formal-index order in graphcupca.cpp: the vector / matrix
// arguments first, the parameter struct last.  Any launch geometry (the graph uses 16 x 16 blocks
// on a ceil(M/16) x ceil(N/16) grid): threads stride over the results.
#pragma once

#include <hip/hip_runtime.h>

// Products and sums round separately, as in pca_cpu: no fused multiply-add, whatever the
// including file's contraction setting.
#pragma clang fp contract(off)

// graphcupca.cpp:163-220 (natural alignment: a double member aligns to 8 on both sides)
struct DCOPY_PARAMS { int N, OFFSETX, OFFSETY, INCX, INCY; };
struct DAXPY_PARAMS { int N; double ALPHA; int OFFSETX, OFFSETY, INCX, INCY; };
struct DGEMV_PARAMS {
    bool TRANSPOSE;
    int M, N;
    double ALPHA, BETA;
    int LDA, OFFSETX, OFFSETY, INCX, INCY;
};
struct DNRM2_PARAMS { int N, OFFSETX, OFFSETY, INCX; };
struct DSCAL_PARAMS { int N; double ALPHA_CONST; int INCX, OFFSETX, OFFSETALPHA_VAR; bool INVERT_ALPHA_VAR; };
struct DGER_PARAMS { int OFFSETY, INCX, INCY, LDA, M, N, OFFSETALPHA_VAR, OFFSETX; };

namespace ptk_pca_blas {

__device__ inline long long gtid() {
    const long long bx = (long long)blockIdx.y * gridDim.x + blockIdx.x;
    return bx * (blockDim.x * blockDim.y * blockDim.z) + (threadIdx.z * blockDim.y + threadIdx.y) * blockDim.x + threadIdx.x;
}
__device__ inline long long gthreads() {
    return (long long)gridDim.x * gridDim.y * gridDim.z * blockDim.x * blockDim.y * blockDim.z;
}

}  // namespace ptk_pca_blas

// y[OFFSETY + i INCY] = x[OFFSETX + i INCX], i < N  (cpublasDcopy, pca_cpu.cpp:147-153)
extern "C" __global__ void ptaskDcopy(const double *X, double *Y, DCOPY_PARAMS p) {
    for (long long i = ptk_pca_blas::gtid(); i < p.N; i += ptk_pca_blas::gthreads())
        Y[p.OFFSETY + i * p.INCY] = X[p.OFFSETX + i * p.INCX];
}

// y += ALPHA x  (cpublasDaxpy, :155-161)
extern "C" __global__ void ptaskDaxpy(const double *X, double *Y, DAXPY_PARAMS p) {
    for (long long i = ptk_pca_blas::gtid(); i < p.N; i += ptk_pca_blas::gthreads())
        Y[p.OFFSETY + i * p.INCY] += p.ALPHA * X[p.OFFSETX + i * p.INCX];
}

// y = ALPHA op(A) x + BETA y, A column-major M x N with leading dimension LDA, x and y offset
// (cpublasDgemv, :163-199: a row's dot product runs in column order from A[row, 0] x[0]).  A and
// y may be one buffer (graphcupca's Gram-Schmidt step): the rows written are never read as A.
extern "C" __global__ void ptaskDgemv(const double *A, const double *X, double *Y, DGEMV_PARAMS p) {
    const double *x = X + p.OFFSETX;
    double *y = Y + p.OFFSETY;
    const int rows = p.TRANSPOSE ? p.N : p.M;
    for (long long row = ptk_pca_blas::gtid(); row < rows; row += ptk_pca_blas::gthreads()) {
        double result = p.BETA * y[row * p.INCY];
        double ax;
        if (!p.TRANSPOSE) {
            ax = A[row] * x[0];
            for (int i = 1; i < p.N; ++i) ax += A[row + (long long)i * p.LDA] * x[(long long)i * p.INCX];
        } else {
            ax = A[row * p.LDA] * x[0];
            for (int i = 1; i < p.M; ++i) ax += A[i + row * p.LDA] * x[(long long)i * p.INCX];
        }
        result += p.ALPHA * ax;
        y[row * p.INCY] = result;
    }
}

// Y[OFFSETY] = sqrt(sum x^2), the sum in index order from 0 (cpublasDnrm2, :253-264)
extern "C" __global__ void ptaskDnrm2(const double *X, double *Y, DNRM2_PARAMS p) {
    if (ptk_pca_blas::gtid() != 0) return;
    double sum = 0.0;
    for (int i = 0; i < p.N; ++i) {
        const double v = X[p.OFFSETX + (long long)i * p.INCX];
        sum += v * v;
    }
    Y[p.OFFSETY] = sqrt(sum);
}

// x *= alpha, alpha = ALPHA_CONST (1 / V[OFFSETALPHA_VAR] or V[OFFSETALPHA_VAR]) (cpublasDscal,
// :201-207, with the scale read from the L vector the graph feeds it)
extern "C" __global__ void ptaskDscal(const double *V, double *X, DSCAL_PARAMS p) {
    const double a = p.INVERT_ALPHA_VAR ? 1.0 / V[p.OFFSETALPHA_VAR] : V[p.OFFSETALPHA_VAR];
    const double alpha = p.ALPHA_CONST * a;
    for (long long i = ptk_pca_blas::gtid(); i < p.N; i += ptk_pca_blas::gthreads()) X[p.OFFSETX + i * p.INCX] *= alpha;
}

// A += alpha x y^T with alpha = -V[OFFSETALPHA_VAR] (the reference fixes ALPHA_CONST at -1,
// graphcupca.cpp:211), x = X + OFFSETX, y = Y + OFFSETY (cpublasDger, :209-251)
extern "C" __global__ void ptaskDger(const double *V, const double *X, const double *Y, double *A, DGER_PARAMS p) {
    const double alpha = -V[p.OFFSETALPHA_VAR];
    const long long cells = (long long)p.M * p.N;
    for (long long c = ptk_pca_blas::gtid(); c < cells; c += ptk_pca_blas::gthreads()) {
        const long long row = c % p.M, col = c / p.M;
        A[col * p.LDA + row] += alpha * X[p.OFFSETX + row * p.INCX] * Y[p.OFFSETY + col * p.INCY];
    }
}
