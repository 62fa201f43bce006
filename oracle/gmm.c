/*
 * oracle/gmm.c -- binary64 restatement of the EM recurrence for a Gaussian mixture.  TEST
 * INFRASTRUCTURE ONLY (the checker; no product code links it).
 *
 * Follows the algorithm of ptaskUnitTest/nr3/gaumixmod.h (Numerical Recipes 3rd ed. Gaumixmod;
 * restated here from the mathematics, not copied) as driven by CPU_EM
 * (ptaskUnitTest/graphemgmm.cpp:92-122):
 *   init (gaumixmod.h:14-27): frac_k = 1/K, Sigma_k = 1e-10 I, then one E-step and one M-step;
 *   E-step (:28-53): r_nk = -1/2 (|L_k^-1 (x_n - mu_k)|^2 + log det Sigma_k) + log frac_k,
 *     then per point tmp_n = max_k r + log sum_k exp(r_nk - max), r_nk = exp(r_nk - tmp_n);
 *     loglike = sum_n tmp_n  (no -D/2 log 2pi term);
 *   M-step (:54-73): w_k = sum_n r_nk, frac_k = w_k/N, mu_k = sum_n r_nk x_n / w_k,
 *     Sigma_k = sum_n r_nk (x_n - mu_k)(x_n - mu_k)^T / w_k with the new mu_k.
 * Covariances are diagonal here (the workload's model); for D = 1 that is the full model of the
 * reference's 1-D tests.  Cholesky of a diagonal matrix is the square root of the diagonal, so
 * |L^-1 u|^2 = sum_m u_m^2 / sigma2_m and log det = sum_m log sigma2_m.
 *
 * A component whose weight is exactly zero keeps its previous mean and variance (the reference
 * would divide by zero); its fraction becomes 0.
 *
 * Pinned by tests/test_oracle_gmm.py: scikit-learn's GaussianMixture run through the same
 * recurrence (same init, diag covariances, no regularisation); K=1 gives the sample mean and
 * variance after the first M-step; log-likelihood non-decreasing; the reference's dist_*.csv
 * fixtures recover their generating parameters.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ptk_oracle.h"

static double gmm_estep(const double *X, long long N, int D, int K, const double *mu,
                        const double *var, const double *frac, double *resp) {
    double *lndet = (double *)malloc(sizeof(double) * (size_t)K);
    for (int k = 0; k < K; ++k) {
        double s = 0;
        for (int m = 0; m < D; ++m) s += log(var[(size_t)k * D + m]);
        lndet[k] = s;
    }
    double loglike = 0;
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
        loglike += tmp;
    }
    free(lndet);
    return loglike;
}

static void gmm_mstep(const double *X, long long N, int D, int K, double *mu, double *var,
                      double *frac, const double *resp) {
    for (int k = 0; k < K; ++k) {
        double w = 0;
        for (long long n = 0; n < N; ++n) w += resp[(size_t)n * K + k];
        frac[k] = w / (double)N;
        if (w == 0) continue;
        for (int m = 0; m < D; ++m) {
            double s = 0;
            for (long long n = 0; n < N; ++n) s += resp[(size_t)n * K + k] * X[(size_t)n * D + m];
            mu[(size_t)k * D + m] = s / w;
        }
        for (int m = 0; m < D; ++m) {
            double s = 0, muk = mu[(size_t)k * D + m];
            for (long long n = 0; n < N; ++n) {
                double u = X[(size_t)n * D + m] - muk;
                s += resp[(size_t)n * K + k] * u * u;
            }
            var[(size_t)k * D + m] = s / w;
        }
    }
}

/* EM as CPU_EM runs it: the constructor's E+M step, then `iters` more.  mu (K*D) holds the
 * initial means on entry and the final means on exit; var (K*D) and frac (K) are outputs;
 * loglike receives iters+1 values (the constructor's E-step first). */
void ptk_oracle_gmm_em(const double *X, long long N, int D, int K, int iters, double *mu,
                       double *var, double *frac, double *loglike) {
    double *resp = (double *)malloc(sizeof(double) * (size_t)N * (size_t)K);
    for (int k = 0; k < K; ++k) {
        frac[k] = 1.0 / K;
        for (int m = 0; m < D; ++m) var[(size_t)k * D + m] = 1.0e-10;
    }
    for (int it = 0; it <= iters; ++it) {
        loglike[it] = gmm_estep(X, N, D, K, mu, var, frac, resp);
        gmm_mstep(X, N, D, K, mu, var, frac, resp);
    }
    free(resp);
}

/* ---- host-task ABI forms of the two GMM task kernels (void f(UINT, void**)) ------------------
 * Same model and partials layout as the device kernels (include/ptk_gmm.h), in binary64:
 * GmmAccumulate(const float* X, GMM_DIMS* d, const double* model, double* partials): E-step from
 * the binary64 mu/var/frac and the moments about the table's binary32 means c, summed over all
 * points into row 0 of partials (rows 1.. are zeroed).
 * GmmFinalize(double* model, const double* partials, GMM_DIMS* d): mu = c + su/w,
 * var = suu/w - (su/w)^2, frac = w/N, the log-likelihood, and the rewritten table. */
void GmmAccumulate(unsigned int n, void **a) {
    (void)n;
    const float *X = (const float *)a[0];
    const GMM_DIMS *d = (const GMM_DIMS *)a[1];
    const double *model = (const double *)a[2];
    double *part = (double *)a[3];
    const int D = d->D, K = d->K, Dp = ptk_gmm_dpad(d);
    const long long S = ptk_gmm_partial_stride(d);
    const double *mu = model + ptk_gmm_mu(d), *var = model + ptk_gmm_var(d), *frac = model + ptk_gmm_frac(d);
    const float *tab = (const float *)(model + ptk_gmm_table(d));
    memset(part, 0, sizeof(double) * (size_t)S * (size_t)d->blocks);
    double *w = part, *su = part + K, *suu = part + K + (size_t)K * D;
    double *lndet = (double *)malloc(sizeof(double) * (size_t)K);
    double *r = (double *)malloc(sizeof(double) * (size_t)K);
    double *x = (double *)malloc(sizeof(double) * (size_t)D);
    for (int k = 0; k < K; ++k) {
        double s = 0;
        for (int m = 0; m < D; ++m) s += log(var[(size_t)k * D + m]);
        lndet[k] = s;
    }
    double loglike = 0;
    for (long long p = 0; p < d->N; ++p) {
        for (int m = 0; m < D; ++m) x[m] = (double)X[(size_t)p * D + m];
        double mx = -99.9e99;
        for (int k = 0; k < K; ++k) {
            double s = 0;
            for (int m = 0; m < D; ++m) {
                double u = x[m] - mu[(size_t)k * D + m];
                s += u * u / var[(size_t)k * D + m];
            }
            r[k] = -0.5 * (s + lndet[k]) + log(frac[k]);
            if (r[k] > mx) mx = r[k];
        }
        double s = 0;
        for (int k = 0; k < K; ++k) s += exp(r[k] - mx);
        const double tmp = mx + log(s);
        loglike += tmp;
        for (int k = 0; k < K; ++k) {
            const double rk = exp(r[k] - tmp);
            w[k] += rk;
            for (int m = 0; m < D; ++m) {
                const double u = x[m] - (double)tab[2 * (size_t)k * Dp + m];
                su[(size_t)k * D + m] += rk * u;
                suu[(size_t)k * D + m] += rk * u * u;
            }
        }
    }
    part[K + 2 * (size_t)K * D] = loglike;
    free(lndet);
    free(r);
    free(x);
}

void GmmFinalize(unsigned int n, void **a) {
    (void)n;
    double *model = (double *)a[0];
    const double *part = (const double *)a[1];
    const GMM_DIMS *d = (const GMM_DIMS *)a[2];
    const int D = d->D, K = d->K, Dp = ptk_gmm_dpad(d);
    const long long S = ptk_gmm_partial_stride(d);
    double *mu = model + ptk_gmm_mu(d), *var = model + ptk_gmm_var(d), *frac = model + ptk_gmm_frac(d);
    const float *tab = (const float *)(model + ptk_gmm_table(d));
    const long long step = (long long)model[0];
    for (int k = 0; k < K; ++k) {
        double w = 0;
        for (int b = 0; b < d->blocks; ++b) w += part[b * S + k];
        frac[k] = w / (double)d->N;
        if (w == 0) continue;
        for (int m = 0; m < D; ++m) {
            const size_t i = (size_t)k * D + m;
            double su = 0, suu = 0;
            for (int b = 0; b < d->blocks; ++b) {
                su += part[b * S + K + i];
                suu += part[b * S + K + (size_t)K * D + i];
            }
            const double dm = su / w;
            mu[i] = (double)tab[2 * (size_t)k * Dp + m] + dm;
            var[i] = suu / w - dm * dm;
        }
    }
    double ll = 0;
    for (int b = 0; b < d->blocks; ++b) ll += part[b * S + K + 2 * (size_t)K * D];
    if (step >= 0 && step <= d->maxIter) model[ptk_gmm_ll(d) + step] = ll;
    model[0] = (double)(step + 1);
    ptk_gmm_fill_table(d, model);
}
