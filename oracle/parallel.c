/*
 * oracle/parallel.c -- multi-threaded forms of the binary64 EM and Gram-Schmidt PCA oracles for
 * the workload-size checks (N = 2,000,000 points; a 65536 x 2048 matrix).  TEST INFRASTRUCTURE
 * ONLY (the checker; no product code links it).
 *
 * Each form performs exactly the floating-point operations of its serial counterpart
 * (oracle/gmm.c ptk_oracle_gmm_em, oracle/pca.c ptk_oracle_pca_gs), in the same order for every
 * accumulated quantity: work is split only across independent results (points, components,
 * matrix rows or columns), never inside one sum, and every sum over points or over matrix
 * columns still runs in increasing index order.  The results are therefore bitwise equal to the
 * serial oracle's for any thread count (tests/test_oracle_gmm.py, tests/test_oracle_pca.py check
 * this).  Loop order differs from the serial code only where the serial loop nest walks memory
 * with a large stride, which changes speed, not results.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "ptk_oracle.h"

/* ---- EM (gaumixmod.h:28-73 as restated in oracle/gmm.c) ---------------------------------- */

static double gmm_estep_mt(const double *X, long long N, int D, int K, const double *mu,
                           const double *var, const double *frac, double *resp, double *tmpn) {
    double *lndet = (double *)malloc(sizeof(double) * (size_t)K);
    for (int k = 0; k < K; ++k) {
        double s = 0;
        for (int m = 0; m < D; ++m) s += log(var[(size_t)k * D + m]);
        lndet[k] = s;
    }
#pragma omp parallel for schedule(static)
    for (long long n = 0; n < N; ++n) {
        const double *x = X + (size_t)n * D;
        double *r = resp + (size_t)n * K;
        double mx = -99.9e99;
        for (int k = 0; k < K; ++k) {
            double sum = 0;
            for (int m = 0; m < D; ++m) {
                double u = x[m] - mu[(size_t)k * D + m];
                sum += u * u / var[(size_t)k * D + m];
            }
            r[k] = -0.5 * (sum + lndet[k]) + log(frac[k]);
            if (r[k] > mx) mx = r[k];
        }
        double sum = 0;
        for (int k = 0; k < K; ++k) sum += exp(r[k] - mx);
        double tmp = mx + log(sum);
        for (int k = 0; k < K; ++k) r[k] = exp(r[k] - tmp);
        tmpn[n] = tmp;
    }
    double loglike = 0;  /* the serial order of the point sum */
    for (long long n = 0; n < N; ++n) loglike += tmpn[n];
    free(lndet);
    return loglike;
}

/* Every (k, m) accumulator adds its points in increasing n, as the serial M-step does; the
 * components are dealt to threads, the points walked in tiles that stay in cache. */
static void gmm_mstep_mt(const double *X, long long N, int D, int K, double *mu, double *var,
                         double *frac, const double *resp) {
    const long long TILE = 2048;
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < K; ++k) {
        double w = 0;
        double *s = (double *)calloc((size_t)D, sizeof(double));
        for (long long n = 0; n < N; ++n) w += resp[(size_t)n * K + k];
        frac[k] = w / (double)N;
        if (w != 0) {
            for (long long n0 = 0; n0 < N; n0 += TILE) {
                const long long n1 = n0 + TILE < N ? n0 + TILE : N;
                for (long long n = n0; n < n1; ++n) {
                    const double r = resp[(size_t)n * K + k];
                    for (int m = 0; m < D; ++m) s[m] += r * X[(size_t)n * D + m];
                }
            }
            for (int m = 0; m < D; ++m) {
                mu[(size_t)k * D + m] = s[m] / w;
                s[m] = 0;
            }
            for (long long n0 = 0; n0 < N; n0 += TILE) {
                const long long n1 = n0 + TILE < N ? n0 + TILE : N;
                for (long long n = n0; n < n1; ++n) {
                    const double r = resp[(size_t)n * K + k];
                    for (int m = 0; m < D; ++m) {
                        double u = X[(size_t)n * D + m] - mu[(size_t)k * D + m];
                        s[m] += r * u * u;
                    }
                }
            }
            for (int m = 0; m < D; ++m) var[(size_t)k * D + m] = s[m] / w;
        }
        free(s);
    }
}

void ptk_oracle_gmm_em_mt(const double *X, long long N, int D, int K, int iters, double *mu,
                          double *var, double *frac, double *loglike, int threads) {
    if (threads > 0) omp_set_num_threads(threads);
    double *resp = (double *)malloc(sizeof(double) * (size_t)N * (size_t)K);
    double *tmpn = (double *)malloc(sizeof(double) * (size_t)N);
    for (int k = 0; k < K; ++k) {
        frac[k] = 1.0 / K;
        for (int m = 0; m < D; ++m) var[(size_t)k * D + m] = 1.0e-10;
    }
    for (int it = 0; it <= iters; ++it) {
        loglike[it] = gmm_estep_mt(X, N, D, K, mu, var, frac, resp, tmpn);
        gmm_mstep_mt(X, N, D, K, mu, var, frac, resp);
    }
    free(resp);
    free(tmpn);
}

/* ---- Gram-Schmidt PCA (pca_cpu.cpp:267-424 as restated in oracle/pca.c) ------------------- */

/* Rows of a column-major M x N matrix are cut into blocks; inside a block every row's sum still
 * walks the columns in increasing n. */
#define PCA_RB 512

void ptk_oracle_pca_gs_mt(const double *X, int M, int N, int K, int J, int center, double *T,
                          double *P, double *R, double *L, int threads) {
    if (threads > 0) omp_set_num_threads(threads);
    const size_t MN = (size_t)M * N;
#pragma omp parallel for schedule(static)
    for (long long n = 0; n < N; ++n) memcpy(R + (size_t)n * M, X + (size_t)n * M, sizeof(double) * (size_t)M);
    memset(T, 0, sizeof(double) * (size_t)M * K);
    memset(P, 0, sizeof(double) * (size_t)N * K);
    (void)MN;
    double *U = (double *)calloc((size_t)(M > K ? M : K), sizeof(double));
    if (center) {
#pragma omp parallel for schedule(static)
        for (int m0 = 0; m0 < M; m0 += PCA_RB) {
            const int m1 = m0 + PCA_RB < M ? m0 + PCA_RB : M;
            for (int m = m0; m < m1; ++m) U[m] = R[m];
            for (int n = 1; n < N; ++n)
                for (int m = m0; m < m1; ++m) U[m] += R[(size_t)n * M + m];
        }
#pragma omp parallel for schedule(static)
        for (int n = 0; n < N; ++n)
            for (int m = 0; m < M; ++m) R[(size_t)n * M + m] += (-1.0 / N) * U[m];
    }
    for (int k = 0; k < K; ++k) {
        double *t = T + (size_t)k * M, *p = P + (size_t)k * N;
        memcpy(t, R + (size_t)k * M, sizeof(double) * (size_t)M);
        for (int j = 0; j < J; ++j) {
#pragma omp parallel for schedule(static)
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
#pragma omp parallel for schedule(static)
            for (int m0 = 0; m0 < M; m0 += PCA_RB) {  /* t = R p, column by column per row block */
                const int m1 = m0 + PCA_RB < M ? m0 + PCA_RB : M;
                for (int m = m0; m < m1; ++m) t[m] = R[m] * p[0];
                for (int n = 1; n < N; ++n) {
                    const double pn = p[n];
                    const double *col = R + (size_t)n * M;
                    for (int m = m0; m < m1; ++m) t[m] += col[m] * pn;
                }
            }
            double lk = 0;
            for (int m = 0; m < M; ++m) lk += t[m] * t[m];
            L[k] = sqrt(lk);
            for (int m = 0; m < M; ++m) t[m] *= 1.0 / L[k];
        }
        const double lk = L[k];
#pragma omp parallel for schedule(static)
        for (int n = 0; n < N; ++n)                  /* R -= L_k t p^T */
            for (int m = 0; m < M; ++m) R[(size_t)n * M + m] += -lk * t[m] * p[n];
    }
    for (int k = 0; k < K; ++k)
        for (int m = 0; m < M; ++m) T[(size_t)k * M + m] *= L[k];
    free(U);
}
