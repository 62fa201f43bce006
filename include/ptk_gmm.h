/*
 * ptk_gmm.h -- the EM-GMM task constants block and the layout of its device-resident model.
 *
 * The reference's EM graph is a stub (ptaskUnitTest/graphemgmm.cpp:124-273); the recurrence is
 * NR3's Gaumixmod (ptaskUnitTest/nr3/gaumixmod.h:14-73) with diagonal covariances.  These
 * layouts are the build's own.
 *
 * GMM_DIMS travels by value (a PTPARM_BYVALSTRUCT sticky port) to both task kernels:
 *   GmmAccumulate(const float* X, GMM_DIMS d, const double* model, double* partials)
 *   GmmFinalize(double* model, const double* partials, GMM_DIMS d)      [model in/out]
 * X is N x D binary32, point-major.  K <= 64, D <= 32.
 */
#ifndef PTK_GMM_H
#define PTK_GMM_H

#if defined(__HIPCC__)
#define PTK_GMM_FN static inline __host__ __device__
#else
#define PTK_GMM_FN static inline
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define PTK_GMM_MAX_K 64
#define PTK_GMM_MAX_D 32

typedef struct _gmm_dims_t {
    int N;        /* points */
    int D;        /* dimensions */
    int K;        /* components */
    int blocks;   /* blocks of GmmAccumulate = rows of the partials buffer */
    int maxIter;  /* capacity of the log-likelihood history */
    int pad[3];
} GMM_DIMS;

/* The model block, binary64 elements:
 *   [0] EM steps done, [1] N, [2] D, [3] K,
 *   mu[K*D], var[K*D], frac[K], loglike[maxIter+1] (one entry per E-step, before its M-step),
 *   then, from an even element, the E-step table in binary32: for each component k, the means
 *   rounded to binary32 and the reciprocal variances, each padded to Dp = D rounded up to 4
 *   (padding 0), followed by C[K] = -1/2 sum_m log var_km + log frac_k.
 * The table is what GmmAccumulate reads; GmmFinalize rewrites it after every M-step and
 * ptk_gmm_fill_table builds it for an initial model.  The binary32 means of the table are also
 * the centres about which the first and second moments are accumulated. */
PTK_GMM_FN int ptk_gmm_dpad(const GMM_DIMS *d) { return (d->D + 3) & ~3; }
PTK_GMM_FN long long ptk_gmm_mu(const GMM_DIMS *d) { (void)d; return 4; }
PTK_GMM_FN long long ptk_gmm_var(const GMM_DIMS *d) { return 4 + (long long)d->K * d->D; }
PTK_GMM_FN long long ptk_gmm_frac(const GMM_DIMS *d) { return 4 + 2ll * d->K * d->D; }
PTK_GMM_FN long long ptk_gmm_ll(const GMM_DIMS *d) { return 4 + 2ll * d->K * d->D + d->K; }
PTK_GMM_FN long long ptk_gmm_table(const GMM_DIMS *d) { return (ptk_gmm_ll(d) + d->maxIter + 2) & ~1ll; }
PTK_GMM_FN long long ptk_gmm_table_floats(const GMM_DIMS *d) { return 2ll * d->K * ptk_gmm_dpad(d) + d->K; }
PTK_GMM_FN long long ptk_gmm_model_elems(const GMM_DIMS *d) { return ptk_gmm_table(d) + (ptk_gmm_table_floats(d) + 1) / 2; }

/* partials (binary64), one row per GmmAccumulate block: w[K], su[K*D], suu[K*D], loglike, with
 * su = sum r (x - c), suu = sum r (x - c)^2 about the table centres c. */
PTK_GMM_FN long long ptk_gmm_partial_stride(const GMM_DIMS *d) { return (long long)d->K + 2ll * d->K * d->D + 1; }

#ifdef __cplusplus
}
#endif

#if !defined(__HIP_DEVICE_COMPILE__)
#include <math.h>
/* Builds the binary32 E-step table of a model from its mu, var and frac (host side). */
static inline void ptk_gmm_fill_table(const GMM_DIMS *d, double *model) {
    const int D = d->D, K = d->K, Dp = ptk_gmm_dpad(d);
    float *tab = (float *)(model + ptk_gmm_table(d));
    const double *mu = model + ptk_gmm_mu(d), *var = model + ptk_gmm_var(d), *frac = model + ptk_gmm_frac(d);
    for (int k = 0; k < K; ++k) {
        double lndet = 0;
        for (int m = 0; m < Dp; ++m) {
            const int i = k * D + m;
            tab[2 * k * Dp + m] = m < D ? (float)mu[i] : 0.0f;
            tab[2 * k * Dp + Dp + m] = m < D ? (float)(1.0 / var[i]) : 0.0f;
            if (m < D) lndet += log(var[i]);
        }
        tab[2 * K * Dp + k] = (float)(-0.5 * lndet + log(frac[k]));
    }
}
#endif

#endif /* PTK_GMM_H */
