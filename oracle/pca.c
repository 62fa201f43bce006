/*
 * oracle/pca.c -- binary64 restatement of the reference's Gram-Schmidt PCA.  TEST INFRASTRUCTURE
 * ONLY (the checker; no product code links it).
 *
 * Follows pca_cpu (ptaskUnitTest/pca_cpu.cpp:267-424) with its BLAS helpers (:147-265) and
 * row_normalize (:60-92), restated from the mathematics:
 *   R = X (column-major M x N); optionally every row has its mean removed (:315-335);
 *   for k < K:  t = R[:,k];
 *               J times: p = R^T t (:353); if k > 0: U = P_<k^T p, p -= P_<k U (:363, :371);
 *                        p *= 1/|p| (:374-376); t = R p (:377); L_k = |t|; t *= 1/L_k (:385-386);
 *               R -= L_k t p^T (:401);
 *   finally T[:,k] *= L_k (:404-408).
 * No convergence test (the reference's is commented out, :387-393).  Outputs are column-major:
 * T (M x K), P (N x K), R (M x N); L (K) is the singular-value estimate sigma_k.
 *
 * Pinned by tests/test_oracle_pca.py: numpy's SVD at convergence on matrices with a known,
 * well-separated spectrum; the exact identity X_centred = T P^T + R; orthonormal P; and the
 * one-iteration case computed independently with numpy.  The BLAS host-task forms at the end of
 * this file are pinned by tests/test_pca_blas_graph.py: the reference's unrolled graph of them
 * must equal ptk_oracle_pca_gs bit for bit, with and without centring.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ptk_oracle.h"

void ptk_oracle_pca_gs(const double *X, int M, int N, int K, int J, int center, double *T, double *P,
                       double *R, double *L) {
    const size_t MN = (size_t)M * N;
    memcpy(R, X, sizeof(double) * MN);
    memset(T, 0, sizeof(double) * (size_t)M * K);
    memset(P, 0, sizeof(double) * (size_t)N * K);
    double *U = (double *)calloc((size_t)(M > K ? M : K), sizeof(double));
    if (center) {
        for (int m = 0; m < M; ++m) U[m] = R[m];
        for (int n = 1; n < N; ++n)
            for (int m = 0; m < M; ++m) U[m] += R[(size_t)n * M + m];
        for (int n = 0; n < N; ++n)
            for (int m = 0; m < M; ++m) R[(size_t)n * M + m] += (-1.0 / N) * U[m];
    }
    for (int k = 0; k < K; ++k) {
        double *t = T + (size_t)k * M, *p = P + (size_t)k * N;
        memcpy(t, R + (size_t)k * M, sizeof(double) * (size_t)M);
        for (int j = 0; j < J; ++j) {
            for (int n = 0; n < N; ++n) {          /* p = R^T t */
                const double *col = R + (size_t)n * M;
                double s = col[0] * t[0];
                for (int m = 1; m < M; ++m) s += col[m] * t[m];
                p[n] = s;
            }
            if (k > 0) {                             /* classical Gram-Schmidt */
                for (int i = 0; i < k; ++i) {
                    const double *q = P + (size_t)i * N;
                    double s = q[0] * p[0];
                    for (int n = 1; n < N; ++n) s += q[n] * p[n];
                    U[i] = s;
                }
                for (int n = 0; n < N; ++n) {
                    double s = P[n] * U[0];
                    for (int i = 1; i < k; ++i) s += P[(size_t)i * N + n] * U[i];
                    p[n] = p[n] + (-1.0 * s);
                }
            }
            double nrm = 0;
            for (int n = 0; n < N; ++n) nrm += p[n] * p[n];
            nrm = sqrt(nrm);
            for (int n = 0; n < N; ++n) p[n] *= 1.0 / nrm;
            for (int m = 0; m < M; ++m) {          /* t = R p */
                double s = R[m] * p[0];
                for (int n = 1; n < N; ++n) s += R[(size_t)n * M + m] * p[n];
                t[m] = s;
            }
            double lk = 0;
            for (int m = 0; m < M; ++m) lk += t[m] * t[m];
            L[k] = sqrt(lk);
            for (int m = 0; m < M; ++m) t[m] *= 1.0 / L[k];
        }
        for (int m = 0; m < M; ++m)                  /* R -= L_k t p^T */
            for (int n = 0; n < N; ++n) R[(size_t)n * M + m] += -L[k] * t[m] * p[n];
    }
    for (int k = 0; k < K; ++k)
        for (int m = 0; m < M; ++m) T[(size_t)k * M + m] *= L[k];
    free(U);
}

/* ---- host-task ABI forms of the PCA task kernels (void f(UINT, void**)) ----------------------
 * Same buffers and layouts as the device kernels (include/ptk_pca.h), binary64 arithmetic, one
 * "block": the partial rows other than row 0 are zero.  By-value PCA_DIMS arrives by pointer. */
#define PCA_L(S, d) ((double *)((char *)(S) + ptk_pca_off_L(d)))
#define PCA_TN(S, d) ((double *)((char *)(S) + ptk_pca_off_tn(d)))
#define PCA_RED(S, d) ((double *)((char *)(S) + ptk_pca_off_red(d)))
#define PCA_P(S, d) ((float *)((char *)(S) + ptk_pca_off_P(d)))
#define PCA_PART(S, d) ((float *)((char *)(S) + ptk_pca_off_part(d)))

void PcaLoad(unsigned int n, void **a) {
    (void)n;
    const float *X = (const float *)a[0];
    const PCA_DIMS *d = (const PCA_DIMS *)a[1];
    float *R = (float *)a[2], *T = (float *)a[3], *S = (float *)a[4];
    memset(R, 0, sizeof(float) * (size_t)d->Mp * d->ld);
    memset(T, 0, sizeof(float) * (size_t)d->Mp * d->K);
    memset(S, 0, (size_t)ptk_pca_state_bytes(d));
    for (int m = 0; m < d->M; ++m) {
        double mean = 0;
        if (d->center) {
            for (int c = 0; c < d->N; ++c) mean += (double)X[(size_t)c * d->M + m];
            mean /= d->N;
        }
        for (int c = 0; c < d->N; ++c) R[(size_t)m * d->ld + c] = (float)((double)X[(size_t)c * d->M + m] - mean);
    }
}

/* t[m] = sum_n R[m][n] v[n] for the rows of R */
static void pca_host_rowdot(const float *R, const PCA_DIMS *d, const float *v, double *t) {
    for (int m = 0; m < d->Mp; ++m) {
        double s = 0;
        for (int c = 0; c < d->ld; ++c) s += (double)R[(size_t)m * d->ld + c] * v[c];
        t[m] = s;
    }
}

/* part row 0 = R^T t, other rows zero */
static void pca_host_coldot(const float *R, const PCA_DIMS *d, const double *t, float *S) {
    float *part = PCA_PART(S, d);
    memset(part, 0, sizeof(float) * (size_t)d->G * d->ld);
    for (int c = 0; c < d->ld; ++c) {
        double s = 0;
        for (int m = 0; m < d->Mp; ++m) s += (double)R[(size_t)m * d->ld + c] * t[m];
        part[c] = (float)s;
    }
}

/* t = R p_c into T[:, c] (binary32), |t|^2 into tn[0] (the other G-1 partials zero), then
 * R -= t p_c^T: the reference's t = R p, L = |t|, t /= L, R -= L t p^T (pca_cpu.cpp:380-401) with
 * L t formed as t itself */
static void pca_host_score_deflate(float *R, float *T, float *S, const PCA_DIMS *d, int c) {
    const float *q = PCA_P(S, d) + (size_t)c * d->ld;
    double *t = (double *)malloc(sizeof(double) * (size_t)d->Mp);
    pca_host_rowdot(R, d, q, t);
    double *tn = PCA_TN(S, d);
    memset(tn, 0, sizeof(double) * (size_t)d->G);
    for (int m = 0; m < d->Mp; ++m) {
        const float tm = (float)t[m];
        T[(size_t)c * d->Mp + m] = tm;
        tn[0] += (double)tm * (double)tm;
        for (int col = 0; col < d->ld; ++col) {
            float *r = R + (size_t)m * d->ld + col;
            *r = (float)((double)*r - (double)tm * q[col]);
        }
    }
    free(t);
}

void PcaStart(unsigned int n, void **a) {
    (void)n;
    float *R = (float *)a[0], *T = (float *)a[1], *S = (float *)a[2];
    const PCA_DIMS *d = (const PCA_DIMS *)a[3];
    const int k = d->k;
    if (k > 0) pca_host_score_deflate(R, T, S, d, k - 1);
    double *t = (double *)malloc(sizeof(double) * (size_t)d->Mp);
    for (int m = 0; m < d->Mp; ++m) t[m] = R[(size_t)m * d->ld + k];
    pca_host_coldot(R, d, t, S);
    free(t);
}

void PcaDeflate(unsigned int n, void **a) {
    (void)n;
    const PCA_DIMS *d = (const PCA_DIMS *)a[3];
    pca_host_score_deflate((float *)a[0], (float *)a[1], (float *)a[2], d, d->k);
}

void PcaGram(unsigned int n, void **a) {
    (void)n;
    float *R = (float *)a[0], *S = (float *)a[1];
    const PCA_DIMS *d = (const PCA_DIMS *)a[2];
    double *t = (double *)malloc(sizeof(double) * (size_t)d->Mp);
    pca_host_rowdot(R, d, PCA_P(S, d) + (size_t)d->k * d->ld, t);
    pca_host_coldot(R, d, t, S);
    free(t);
}

void PcaReduce(unsigned int n, void **a) {
    (void)n;
    float *S = (float *)a[0];
    const PCA_DIMS *d = (const PCA_DIMS *)a[1];
    const float *part = PCA_PART(S, d);
    double *red = PCA_RED(S, d);
    for (int g = 0; g < d->GR; ++g) {
        const int lo = (int)((long long)g * d->G / d->GR), hi = (int)((long long)(g + 1) * d->G / d->GR);
        for (int c = 0; c < d->ld; ++c) {
            double s = 0;
            for (int b = lo; b < hi; ++b) s += part[(size_t)b * d->ld + c];
            red[(size_t)g * d->ld + c] = s;
        }
    }
}

void PcaOrthonormalize(unsigned int n, void **a) {
    (void)n;
    float *S = (float *)a[0];
    const PCA_DIMS *d = (const PCA_DIMS *)a[1];
    const double *red = PCA_RED(S, d);
    float *P = PCA_P(S, d);
    const int k = d->k, ld = d->ld;
    double *p = (double *)calloc((size_t)ld, sizeof(double));
    double *u = (double *)calloc((size_t)(k > 0 ? k : 1), sizeof(double));
    for (int c = 0; c < ld; ++c)
        for (int g = 0; g < d->GR; ++g) p[c] += red[(size_t)g * ld + c];
    for (int j = 0; j < k; ++j)
        for (int c = 0; c < ld; ++c) u[j] += (double)P[(size_t)j * ld + c] * p[c];
    for (int c = 0; c < ld; ++c) {
        double s = 0;
        for (int j = 0; j < k; ++j) s += (double)P[(size_t)j * ld + c] * u[j];
        p[c] -= s;
    }
    double nrm = 0;
    for (int c = 0; c < ld; ++c) nrm += p[c] * p[c];
    const double inv = 1.0 / sqrt(nrm);
    for (int c = 0; c < ld; ++c) P[(size_t)k * ld + c] = (float)(p[c] * inv);
    free(p);
    free(u);
}

void PcaFinish(unsigned int n, void **a) {
    (void)n;
    float *S = (float *)a[0];
    const PCA_DIMS *d = (const PCA_DIMS *)a[1];
    double s = 0;
    for (int b = 0; b < d->G; ++b) s += PCA_TN(S, d)[b];
    PCA_L(S, d)[d->k] = sqrt(s);
}

/* ---- host-task forms of the reference PCA graph's BLAS tasks (ptaskUnitTest/graphcupca.cpp:294-299)
 * with the by-value parameter structs of graphcupca.cpp:161-220 (passed by pointer to a host
 * task); the same operations as pca_cpu's cpublasD* routines (pca_cpu.cpp:147-265). */
typedef struct { int N, OFFSETX, OFFSETY, INCX, INCY; } PTK_DCOPY;
typedef struct { int N; double ALPHA; int OFFSETX, OFFSETY, INCX, INCY; } PTK_DAXPY;
typedef struct { _Bool TRANSPOSE; int M, N; double ALPHA, BETA; int LDA, OFFSETX, OFFSETY, INCX, INCY; } PTK_DGEMV;
typedef struct { int N, OFFSETX, OFFSETY, INCX; } PTK_DNRM2;
typedef struct { int N; double ALPHA_CONST; int INCX, OFFSETX, OFFSETALPHA_VAR; _Bool INVERT_ALPHA_VAR; } PTK_DSCAL;
typedef struct { int OFFSETY, INCX, INCY, LDA, M, N, OFFSETALPHA_VAR, OFFSETX; } PTK_DGER;

void ptaskDcopy(unsigned int n, void **a) {
    (void)n;
    const double *X = (const double *)a[0];
    double *Y = (double *)a[1];
    const PTK_DCOPY *p = (const PTK_DCOPY *)a[2];
    for (long long i = 0; i < p->N; ++i) Y[p->OFFSETY + i * p->INCY] = X[p->OFFSETX + i * p->INCX];
}

void ptaskDaxpy(unsigned int n, void **a) {
    (void)n;
    const double *X = (const double *)a[0];
    double *Y = (double *)a[1];
    const PTK_DAXPY *p = (const PTK_DAXPY *)a[2];
    for (long long i = 0; i < p->N; ++i) Y[p->OFFSETY + i * p->INCY] += p->ALPHA * X[p->OFFSETX + i * p->INCX];
}

void ptaskDgemv(unsigned int n, void **a) {
    (void)n;
    const double *A = (const double *)a[0];
    const PTK_DGEMV *p = (const PTK_DGEMV *)a[3];
    const double *x = (const double *)a[1] + p->OFFSETX;
    double *y = (double *)a[2] + p->OFFSETY;
    const int rows = p->TRANSPOSE ? p->N : p->M;
    for (long long row = 0; row < rows; ++row) {
        double result = p->BETA * y[row * p->INCY], ax;
        if (!p->TRANSPOSE) {
            ax = A[row] * x[0];
            for (int i = 1; i < p->N; ++i) ax += A[row + (long long)i * p->LDA] * x[(long long)i * p->INCX];
        } else {
            ax = A[row * p->LDA] * x[0];
            for (int i = 1; i < p->M; ++i) ax += A[i + row * p->LDA] * x[(long long)i * p->INCX];
        }
        result += p->ALPHA * ax;
        y[row * p->INCY] = result;
    }
}

void ptaskDnrm2(unsigned int n, void **a) {
    (void)n;
    const double *X = (const double *)a[0];
    double *Y = (double *)a[1];
    const PTK_DNRM2 *p = (const PTK_DNRM2 *)a[2];
    double sum = 0.0;
    for (int i = 0; i < p->N; ++i) {
        const double v = X[p->OFFSETX + (long long)i * p->INCX];
        sum += v * v;
    }
    Y[p->OFFSETY] = sqrt(sum);
}

void ptaskDscal(unsigned int n, void **a) {
    (void)n;
    const double *V = (const double *)a[0];
    double *X = (double *)a[1];
    const PTK_DSCAL *p = (const PTK_DSCAL *)a[2];
    const double s = p->INVERT_ALPHA_VAR ? 1.0 / V[p->OFFSETALPHA_VAR] : V[p->OFFSETALPHA_VAR];
    const double alpha = p->ALPHA_CONST * s;
    for (long long i = 0; i < p->N; ++i) X[p->OFFSETX + i * p->INCX] *= alpha;
}

void ptaskDger(unsigned int n, void **a) {
    (void)n;
    const double *V = (const double *)a[0], *X = (const double *)a[1], *Y = (const double *)a[2];
    double *A = (double *)a[3];
    const PTK_DGER *p = (const PTK_DGER *)a[4];
    const double alpha = -V[p->OFFSETALPHA_VAR];
    for (long long col = 0; col < p->N; ++col)
        for (long long row = 0; row < p->M; ++row)
            A[col * p->LDA + row] += alpha * X[p->OFFSETX + row * p->INCX] * Y[p->OFFSETY + col * p->INCY];
}
