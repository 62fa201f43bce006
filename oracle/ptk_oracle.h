/*
 * ptk_oracle.h -- CPU restatements of the reference algorithms on the PTask hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  These functions are the checker: tests, the smoke check and the
 * CPU-baseline leg of the benchmark call them to produce expected results.  Nothing in the
 * product (ptask_amd/, include/) includes, links or loads them.
 */
#ifndef PTK_ORACLE_H
#define PTK_ORACLE_H

#include <stdint.h>

#include "ptk_fdtd.h"
#include "ptk_gmm.h"
#include "ptk_pca.h"
#include "ptk_prng.h"

#ifdef __cplusplus
extern "C" {
#endif

/* FDTD (oracle/fdtd.c) */
void ptk_oracle_fdtd_hx_f32(const float *Hx, const float *Ey, const float *Ez, const FDTD_PARAMS *p, float *o);
void ptk_oracle_fdtd_hy_f32(const float *Hy, const float *Ex, const float *Ez, const FDTD_PARAMS *p, float *o);
void ptk_oracle_fdtd_hz_f32(const float *Hz, const float *Ex, const float *Ey, const FDTD_PARAMS *p, float *o);
void ptk_oracle_fdtd_ex_f32(const float *Ex, const float *Hy, const float *Hz, const FDTD_PARAMS *p, float *o);
void ptk_oracle_fdtd_ey_f32(const float *Ey, const float *Hx, const float *Hz, const FDTD_PARAMS *p, float *o);
void ptk_oracle_fdtd_ez_f32(const float *Ez, const float *Hx, const float *Hy, const FDTD_PARAMS *p, float *o);
void ptk_oracle_fdtd_hx_f64(const double *Hx, const double *Ey, const double *Ez, const FDTD_PARAMS *p, double *o);
void ptk_oracle_fdtd_hy_f64(const double *Hy, const double *Ex, const double *Ez, const FDTD_PARAMS *p, double *o);
void ptk_oracle_fdtd_hz_f64(const double *Hz, const double *Ex, const double *Ey, const FDTD_PARAMS *p, double *o);
void ptk_oracle_fdtd_ex_f64(const double *Ex, const double *Hy, const double *Hz, const FDTD_PARAMS *p, double *o);
void ptk_oracle_fdtd_ey_f64(const double *Ey, const double *Hx, const double *Hz, const FDTD_PARAMS *p, double *o);
void ptk_oracle_fdtd_ez_f64(const double *Ez, const double *Hx, const double *Hy, const FDTD_PARAMS *p, double *o);
void ptk_oracle_fdtd_run_f32(float *f[6], float *scratch[6], const FDTD_PARAMS *p, int steps);
void ptk_oracle_fdtd_run_f64(double *f[6], double *scratch[6], const FDTD_PARAMS *p, int steps);
void ptk_oracle_fdtd_run_ptrs_f32(float **f, float **scratch, const FDTD_PARAMS *p, int steps);
void ptk_oracle_fdtd_run_ptrs_f64(double **f, double **scratch, const FDTD_PARAMS *p, int steps);
void ptk_oracle_fdtd_init_field(float *a, int field, int Nx, int Ny, int Nz, uint64_t seed);

/* Prefix sum and groupby (oracle/groupby.c) */
void ptk_oracle_pfxsum_i32(const int32_t *in, int32_t *out, long long n);
void ptk_oracle_pfxsum_i64(const int64_t *in, int64_t *out, long long n);
void ptk_oracle_groupby(const int32_t *keys, const int32_t *vals, long long n, int numKeys,
                        int64_t *groups, int64_t *offsets, int32_t *outKeys, int32_t *outVals);

/* EM for a diagonal Gaussian mixture (oracle/gmm.c) */
void ptk_oracle_gmm_em(const double *X, long long N, int D, int K, int iters, double *mu,
                       double *var, double *frac, double *loglike);

/* Gram-Schmidt PCA (oracle/pca.c); column-major X (M x N), T (M x K), P (N x K), R (M x N) */
void ptk_oracle_pca_gs(const double *X, int M, int N, int K, int J, int center, double *T, double *P,
                       double *R, double *L);

/* Multi-threaded forms (oracle/parallel.c): bitwise equal to the serial ones for any thread
 * count (threads <= 0: the OpenMP default). */
void ptk_oracle_gmm_em_mt(const double *X, long long N, int D, int K, int iters, double *mu,
                          double *var, double *frac, double *loglike, int threads);
void ptk_oracle_pca_gs_mt(const double *X, int M, int N, int K, int J, int center, double *T,
                          double *P, double *R, double *L, int threads);

/* Host-task ABI forms (void f(UINT, void**)), loaded by name by the HostAccelerator. */
void pfxsum(unsigned int n, void **a);
void GroupbyHistogram(unsigned int n, void **a);
void GroupbyBucketScan(unsigned int n, void **a);
void GroupbyPartition(unsigned int n, void **a);
void GroupbyKeyCount(unsigned int n, void **a);
void GroupbyKeyScan(unsigned int n, void **a);
void GroupbyReshuffle(unsigned int n, void **a);
void GroupbyTotals(unsigned int n, void **a);
void coalesce_int_float(unsigned int n, void **a);
void GroupbyCoalesce(unsigned int n, void **a);
void GroupbyCoalesceSizes(unsigned int n, void **a);
void HxComputation(unsigned int n, void **a);
void HyComputation(unsigned int n, void **a);
void HzComputation(unsigned int n, void **a);
void ExComputation(unsigned int n, void **a);
void EyComputation(unsigned int n, void **a);
void EzComputation(unsigned int n, void **a);
void FdtdHFields(unsigned int n, void **a);
void FdtdEFields(unsigned int n, void **a);
void FdtdStepFused(unsigned int n, void **a);
void FdtdStepOverlap(unsigned int n, void **a);
void FdtdStepOverlap8(unsigned int n, void **a);
void FdtdTwoStepOverlap(unsigned int n, void **a);
void FdtdThreeStepOverlap(unsigned int n, void **a);
void FdtdThreeStepPersistent(unsigned int n, void **a);
void GmmAccumulate(unsigned int n, void **a);
void GmmFinalize(unsigned int n, void **a);
void PcaLoad(unsigned int n, void **a);
void PcaStart(unsigned int n, void **a);
void PcaGram(unsigned int n, void **a);
void PcaDeflate(unsigned int n, void **a);
void PcaReduce(unsigned int n, void **a);
void PcaOrthonormalize(unsigned int n, void **a);
void PcaFinish(unsigned int n, void **a);

#ifdef __cplusplus
}
#endif

#endif /* PTK_ORACLE_H */
