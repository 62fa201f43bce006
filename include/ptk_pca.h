/*
 * ptk_pca.h -- constants block and device layouts of the Gram-Schmidt PCA task kernels.
 *
 * The algorithm is the reference's GS-PCA (ptaskUnitTest/pca_cpu.cpp:267-424, the unrolled BLAS
 * task graph of ptaskUnitTest/graphcupca.cpp:317-1540): for each component k, t = R[:,k]; J times
 * { p = R^T t; p -= P_<k (P_<k^T p); p /= |p|; t = R p; L_k = |t|; t /= L_k }; R -= L_k t p^T;
 * finally T[:,k] *= L_k.  The layouts below are the build's own.
 *
 * Input X: M x N binary32, column-major (the reference's COLUMN_MAJOR_INDEX(m, n, M)).
 * R: the residual on the device, row-major Mp x ld binary32 (ld = N rounded up to 4, Mp = M
 *    rounded up to 64; padding is zero).
 * T: K x Mp binary32 (column k of the scores at T + k*Mp).
 * S: the state block (byte offsets below): L[K] and tn[G] binary64, red[GR][ld] binary64,
 *    P[K][ld] binary32 (the loads, component k at P + k*ld), part[G][ld] binary32.
 *
 * Task kernels (PCA_DIMS by value; k selects the component a task works on):
 *   PcaLoad(const float* X, PCA_DIMS d, float* R, float* T, float* S)
 *   PcaStart(float* R, float* T, float* S, PCA_DIMS d)      for k > 0: t_{k-1} = R p_{k-1} into
 *                                                           T[:,k-1] and the |t|^2 partials,
 *                                                           R -= t p_{k-1}^T; then R^T R[:,k]
 *   PcaGram(float* R, float* S, PCA_DIMS d)                 R^T (R p_k)
 *   PcaDeflate(float* R, float* T, float* S, PCA_DIMS d)    t_k = R p_k into T[:,k] and the |t|^2
 *                                                           partials, R -= t_k p_k^T (the last k)
 *   PcaReduce(float* S, PCA_DIMS d)                         partial sums, first level
 *   PcaOrthonormalize(float* S, PCA_DIMS d)                 p_k = GS(p), p_k /= |p_k|
 *   PcaFinish(float* S, PCA_DIMS d)                         L_k = |t_k| from the partials
 * T[:,k] = t_k = R p_k is the reference's final L_k (t / L_k), formed without the two roundings.
 */
#ifndef PTK_PCA_H
#define PTK_PCA_H

#if defined(__HIPCC__)
#define PTK_PCA_FN static inline __host__ __device__
#else
#define PTK_PCA_FN static inline
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define PTK_PCA_MAX_N 4096 /* columns */
#define PTK_PCA_MAX_K 64   /* components */
#define PTK_PCA_ROWS 64    /* row padding of R and T */

typedef struct _pca_dims_t {
    int M, N, K, J; /* rows, columns, components, inner (power) iterations */
    int ld;         /* row stride of R: N rounded up to 4 */
    int Mp;         /* rows of R and T: M rounded up to 64 */
    int G;          /* blocks of the row passes = rows of part and tn */
    int GR;         /* rows of red (first-level reduction groups), 1 <= GR <= G */
    int k;          /* the component this task works on */
    int center;     /* subtract each row's mean from X when loading (pca_cpu.cpp:315-335) */
    int pad[2];
} PCA_DIMS;

PTK_PCA_FN int ptk_pca_ld(int N) { return (N + 3) & ~3; }
PTK_PCA_FN int ptk_pca_mp(int M) { return (M + PTK_PCA_ROWS - 1) / PTK_PCA_ROWS * PTK_PCA_ROWS; }

/* byte offsets into the state block S */
PTK_PCA_FN long long ptk_pca_off_L(const PCA_DIMS *d) { (void)d; return 0; }
PTK_PCA_FN long long ptk_pca_off_tn(const PCA_DIMS *d) { return 8ll * ((d->K + 1) & ~1); }
PTK_PCA_FN long long ptk_pca_off_red(const PCA_DIMS *d) { return ptk_pca_off_tn(d) + 8ll * ((d->G + 1) & ~1); }
PTK_PCA_FN long long ptk_pca_off_P(const PCA_DIMS *d) { return ptk_pca_off_red(d) + 8ll * d->GR * d->ld; }
PTK_PCA_FN long long ptk_pca_off_part(const PCA_DIMS *d) { return ptk_pca_off_P(d) + 4ll * d->K * d->ld; }
PTK_PCA_FN long long ptk_pca_state_bytes(const PCA_DIMS *d) { return ptk_pca_off_part(d) + 4ll * d->G * d->ld; }

#ifdef __cplusplus
}
#endif

#endif /* PTK_PCA_H */
