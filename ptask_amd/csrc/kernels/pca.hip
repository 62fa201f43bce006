// pca.hip -- Gram-Schmidt PCA (NIPALS with deflation) as PTask task kernels (pca.hsaco).
//
// Semantics follow ptaskUnitTest/pca_cpu.cpp:267-424 (the host reference the unrolled BLAS graph
// of ptaskUnitTest/graphcupca.cpp is checked against); the decomposition into kernels is the
// build's own.  The reference runs, per component k and inner iteration, separate BLAS-2 tasks
// (dgemv R^T t, dgemv for the Gram-Schmidt step, dnrm2, dscal, dgemv R p, dnrm2, dscal) and a
// rank-1 dger per component.  Here the matrix R lives row-major in HBM and every pass over it
// does as much as one read allows:
//   PcaStart    t_{k-1} = R p_{k-1} and deflation by component k-1 (R -= t p^T, written back),
//               fused with p = R^T R[:,k]: the last inner iteration's t = R p of component k-1
//               is formed in the same read of R that deflates it
//   PcaGram     p = R^T (R p) in one read of R: each 8-row slab's t = R_slab p is reduced inside
//               the block and immediately used for R_slab^T t (the reference's t = R p, |t|,
//               t /= |t|, p = R^T t -- the scaling by 1/|t| is undone by the next normalisation)
//   PcaDeflate  t_K-1 = R p and R -= t p^T by the last component (the residual the reference returns)
// T[:,k] holds t = R p_k, which is the reference's final L_k (t / L_k) (pca_cpu.cpp:385-386 and
// 404-408) without the two rounding steps.  The small steps are single kernels: PcaReduce +
// PcaOrthonormalize (sum the per-block p partials in a fixed order, p -= P_<k (P_<k^T p),
// p /= |p|) and PcaFinish (L_k = |t| from the per-block |t|^2 partials).  All reductions are
// fixed-order, so results are bitwise reproducible.  Dot products and norms of length-N and
// length-M vectors are binary64.
#include <hip/hip_runtime.h>

#include "ptk_pca.h"

// Tolerance-checked binary32 work: let the compiler fuse multiply-adds (the build default is off).
#pragma clang fp contract(fast)

enum { PCA_START = 0, PCA_GRAM = 1, PCA_DEFLATE = 3 };

__device__ inline bool pca_ok(const PCA_DIMS &d) {
    return d.M >= 1 && d.N >= 1 && d.N <= PTK_PCA_MAX_N && d.K >= 1 && d.K <= PTK_PCA_MAX_K &&
           d.K <= d.N && d.J >= 1 && d.ld == ptk_pca_ld(d.N) && d.Mp == ptk_pca_mp(d.M) && d.G >= 1 &&
           d.GR >= 1 && d.GR <= d.G && d.k >= 0 && d.k < d.K;
}

__device__ inline double *pca_L(float *S, const PCA_DIMS &d) { return (double *)((char *)S + ptk_pca_off_L(&d)); }
__device__ inline double *pca_tn(float *S, const PCA_DIMS &d) { return (double *)((char *)S + ptk_pca_off_tn(&d)); }
__device__ inline double *pca_red(float *S, const PCA_DIMS &d) { return (double *)((char *)S + ptk_pca_off_red(&d)); }
__device__ inline float *pca_P(float *S, const PCA_DIMS &d) { return (float *)((char *)S + ptk_pca_off_P(&d)); }
__device__ inline float *pca_part(float *S, const PCA_DIMS &d) { return (float *)((char *)S + ptk_pca_off_part(&d)); }

__device__ inline float f4get(const float4 &v, int e) { return e == 0 ? v.x : e == 1 ? v.y : e == 2 ? v.z : v.w; }
__device__ inline float f4dot(const float4 &a, const float4 &b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }

// ---- PcaLoad: X (column-major M x N) -> R (row-major Mp x ld), rows optionally centred; T and S
// cleared.  Grid Mp/64 blocks of 256; block b owns rows 64b .. 64b+63.
extern "C" __global__ void __launch_bounds__(256)
PcaLoad(const float *__restrict__ X, PCA_DIMS d, float *__restrict__ R, float *__restrict__ T,
        float *__restrict__ S) {
    __shared__ double sSum[4][64];
    __shared__ double sMean[64];
    __shared__ float tile[64][65];
    if (!pca_ok(d) || blockDim.x != 256 || (int)gridDim.x != d.Mp / 64) return;
    const int tid = threadIdx.x, M = d.M, N = d.N, ld = d.ld;
    const long long r0 = (long long)blockIdx.x * 64;
    {
        // row sums: eight columns' loads issued together (row and column indexes clamped in
        // range), added in column order
        const int r = tid & 63, grp = tid >> 6;
        const long long m = r0 + r < M ? r0 + r : M - 1;
        double s = 0;
        if (d.center && r0 + r < M)
            for (int n0 = grp; n0 < N; n0 += 32) {
                float v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int n = n0 + 4 * u < N ? n0 + 4 * u : N - 1;
                    v[u] = X[(long long)n * M + m];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (n0 + 4 * u < N) s += (double)v[u];
            }
        sSum[grp][r] = s;
    }
    __syncthreads();
    if (tid < 64) sMean[tid] = d.center ? (sSum[0][tid] + sSum[1][tid] + sSum[2][tid] + sSum[3][tid]) / N : 0.0;
    __syncthreads();
    // 64 x 64 tiles through LDS; the next tile's sixteen loads per thread are issued before the
    // current tile is transposed and stored
    const int r = tid & 63, c0 = tid >> 6;
    const long long m = r0 + r < M ? r0 + r : M - 1;
    const bool mv = r0 + r < M;
    float cur[16], nxt[16];
    auto load = [&](float (&v)[16], int n0) {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int n = n0 + c0 + 4 * j;
            v[j] = X[(long long)(n < N ? n : N - 1) * M + m];
        }
    };
    load(cur, 0);
    for (int n0 = 0; n0 < ld; n0 += 64) {
        if (n0 + 64 < ld) load(nxt, n0 + 64);
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int c = c0 + 4 * j;
            tile[c][r] = (n0 + c < N && mv) ? (float)((double)cur[j] - sMean[r]) : 0.f;
        }
        __syncthreads();
        for (int i = tid; i < 64 * 64; i += 256) {
            const int rr = i >> 6, c = i & 63;
            if (n0 + c < ld) R[(r0 + rr) * ld + n0 + c] = tile[c][rr];
        }
        __syncthreads();
        if (n0 + 64 < ld)
#pragma unroll
            for (int j = 0; j < 16; ++j) cur[j] = nxt[j];
    }
    const long long nT = (long long)d.K * d.Mp, nS = ptk_pca_state_bytes(&d) / 4;
    const long long step = (long long)gridDim.x * 256;
    for (long long i = (long long)blockIdx.x * 256 + tid; i < nT; i += step) T[i] = 0.f;
    for (long long i = (long long)blockIdx.x * 256 + tid; i < nS; i += step) S[i] = 0.f;
}

// ---- the row passes.  Block of 256; thread t owns float4 column chunks t + 256 c (c < NCH); a
// slab of RB rows is held in registers, its RB row dot products are reduced across the block
// (wave shuffles, then a double-buffered LDS exchange, one barrier per slab).
//
// Deflation by component dk (PcaStart of component k deflates k - 1; PcaDeflate deflates the
// last): the rows' scores t = R p_dk are the block-reduced dot products of the rows about to be
// deflated (pca_cpu.cpp:380-401 computes t = R p, L = |t|, t /= L and R -= L t p^T; L t is t
// before the scaling, so no separate pass computes it); t goes to T[:, dk] and its squares to
// the block's tn partial (PcaFinish takes L_dk = |t| from them); then R -= t p_dk^T.
template <int NCH, int RB, int MODE>
__device__ __forceinline__ void pca_row_pass(float *__restrict__ R, float *__restrict__ T,
                                             float *__restrict__ S, const PCA_DIMS &d) {
    __shared__ float sRed[2][4][RB];
    __shared__ float sCol[2][RB];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int ld = d.ld, ld4 = ld >> 2, k = d.k;
    const long long Mp = d.Mp;
    float *P = pca_P(S, d);
    int cix[NCH];
    bool cv[NCH];
    float4 wv4[NCH], acc[NCH];  // the vector the rows are dotted with: p_k (Gram) or p_dk (deflation)
    const bool deflate = (MODE == PCA_START && k > 0) || MODE == PCA_DEFLATE;
    const bool rowdot = MODE == PCA_GRAM || deflate;
    const int dk = MODE == PCA_START ? k - 1 : k;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        cix[c] = tid + 256 * c;
        cv[c] = cix[c] < ld4;
        const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
        wv4[c] = (rowdot && cv[c]) ? ((const float4 *)(P + (long long)(MODE == PCA_GRAM ? k : dk) * ld))[cix[c]] : z;
        acc[c] = z;
    }
    // p_dk at column k: every thread forms column k of the deflated rows from the exchanged
    // pre-deflation values, as the column's owner deflates it (the same fused multiply-add)
    const float pk = (MODE == PCA_START && deflate) ? P[(long long)dk * ld + k] : 0.f;
    const int ck = k >> 2, ce = k & 3, cowner = ck & 255, creg = ck >> 8;
    const long long nslab = Mp / RB;
    int buf = 0;
    double tsq = 0;
    for (long long s = blockIdx.x; s < nslab; s += gridDim.x) {
        const long long r0 = s * RB;
        float4 x[RB][NCH];
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
            for (int c = 0; c < NCH; ++c)
                x[i][c] = cv[c] ? ((const float4 *)(R + (r0 + i) * ld))[cix[c]] : make_float4(0.f, 0.f, 0.f, 0.f);
        if (rowdot) {
            float v[RB];
#pragma unroll
            for (int i = 0; i < RB; ++i) {
                v[i] = 0.f;
#pragma unroll
                for (int c = 0; c < NCH; ++c) v[i] += f4dot(x[i][c], wv4[c]);
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) v[i] += __shfl_xor(v[i], off, 64);
            }
            if (lane == 0) {
#pragma unroll
                for (int i = 0; i < RB; ++i) sRed[buf][wv][i] = v[i];
            }
        }
        if (MODE == PCA_START && tid == cowner) {
            // column k of the rows (before any deflation), from the thread that owns it
#pragma unroll
            for (int i = 0; i < RB; ++i) {
                float xv = 0.f;
#pragma unroll
                for (int c = 0; c < NCH; ++c)
                    if (c == creg) xv = f4get(x[i][c], ce);
                sCol[buf][i] = xv;
            }
        }
        __syncthreads();
        float red[RB];
#pragma unroll
        for (int i = 0; i < RB; ++i)
            red[i] = rowdot ? sRed[buf][0][i] + sRed[buf][1][i] + sRed[buf][2][i] + sRed[buf][3][i] : 0.f;
        if (deflate) {
#pragma unroll
            for (int i = 0; i < RB; ++i) {
                const float a = red[i];  // t of row r0 + i
#pragma unroll
                for (int c = 0; c < NCH; ++c) {
                    if (!cv[c]) continue;
                    x[i][c].x = fmaf(-a, wv4[c].x, x[i][c].x);
                    x[i][c].y = fmaf(-a, wv4[c].y, x[i][c].y);
                    x[i][c].z = fmaf(-a, wv4[c].z, x[i][c].z);
                    x[i][c].w = fmaf(-a, wv4[c].w, x[i][c].w);
                    ((float4 *)(R + (r0 + i) * ld))[cix[c]] = x[i][c];
                }
            }
            if (tid == 0) {
#pragma unroll
                for (int i = 0; i < RB; ++i) {
                    T[(long long)dk * Mp + r0 + i] = red[i];
                    tsq += (double)red[i] * (double)red[i];
                }
            }
        }
        if (MODE == PCA_DEFLATE) {
            buf ^= 1;
            continue;
        }
        float t[RB];  // Start: column k of the (deflated) residual; Gram: R_slab p_k
#pragma unroll
        for (int i = 0; i < RB; ++i)
            t[i] = MODE == PCA_GRAM ? red[i] : (deflate ? fmaf(-red[i], pk, sCol[buf][i]) : sCol[buf][i]);
        buf ^= 1;
#pragma unroll
        for (int c = 0; c < NCH; ++c)
#pragma unroll
            for (int i = 0; i < RB; ++i) {
                acc[c].x = fmaf(x[i][c].x, t[i], acc[c].x);
                acc[c].y = fmaf(x[i][c].y, t[i], acc[c].y);
                acc[c].z = fmaf(x[i][c].z, t[i], acc[c].z);
                acc[c].w = fmaf(x[i][c].w, t[i], acc[c].w);
            }
    }
    if (deflate && tid == 0) pca_tn(S, d)[blockIdx.x] = tsq;
    if (MODE != PCA_DEFLATE) {
        float *part = pca_part(S, d) + (long long)blockIdx.x * ld;
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            if (cv[c]) ((float4 *)part)[cix[c]] = acc[c];
    }
}

template <int MODE>
__device__ __forceinline__ void pca_row_dispatch(float *R, float *T, float *S, const PCA_DIMS &d) {
    switch ((d.ld / 4 + 255) / 256) {
        case 1: pca_row_pass<1, 8, MODE>(R, T, S, d); break;
        case 2: pca_row_pass<2, 8, MODE>(R, T, S, d); break;
        case 3: pca_row_pass<3, 4, MODE>(R, T, S, d); break;
        default: pca_row_pass<4, 4, MODE>(R, T, S, d); break;
    }
}

extern "C" __global__ void __launch_bounds__(256)
PcaStart(float *__restrict__ R, float *__restrict__ T, float *__restrict__ S, PCA_DIMS d) {
    if (!pca_ok(d) || blockDim.x != 256 || (int)gridDim.x != d.G) return;
    pca_row_dispatch<PCA_START>(R, T, S, d);
}

extern "C" __global__ void __launch_bounds__(256)
PcaGram(float *__restrict__ R, float *__restrict__ S, PCA_DIMS d) {
    if (!pca_ok(d) || blockDim.x != 256 || (int)gridDim.x != d.G) return;
    pca_row_dispatch<PCA_GRAM>(R, nullptr, S, d);
}

// After the last component: t_k = R p_k into T[:, k] and the tn partials, R -= t_k p_k^T for
// k = d.k (pca_cpu.cpp:380-401 of the final k).
extern "C" __global__ void __launch_bounds__(256)
PcaDeflate(float *__restrict__ R, float *__restrict__ T, float *__restrict__ S, PCA_DIMS d) {
    if (!pca_ok(d) || blockDim.x != 256 || (int)gridDim.x != d.G) return;
    pca_row_dispatch<PCA_DEFLATE>(R, T, S, d);
}

// ---- PcaReduce: red[g][n] = sum of part[b][n] over the g-th of GR contiguous groups of the G
// rows.  Grid (ceil(ld/64), GR) blocks of 256 (64 columns x 4 interleaved row subsets).  The
// loads read clamped, always valid rows unconditionally, eight at a time, so they issue together;
// the sums take the valid terms in row order.
extern "C" __global__ void __launch_bounds__(256)
PcaReduce(float *__restrict__ S, PCA_DIMS d) {
    __shared__ double sh[4][64];
    if (!pca_ok(d) || blockDim.x != 256 || (int)gridDim.x != (d.ld + 63) / 64 || (int)gridDim.y != d.GR) return;
    const int tid = threadIdx.x, c = tid & 63, sub = tid >> 6, g = blockIdx.y, ld = d.ld;
    const int col = blockIdx.x * 64 + c, cc = col < ld ? col : ld - 1;
    const int lo = (int)((long long)g * d.G / d.GR), hi = (int)((long long)(g + 1) * d.G / d.GR);
    const float *part = pca_part(S, d);
    double s = 0;
    for (int b0 = lo + sub; b0 < hi; b0 += 32) {
        float x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int b = b0 + 4 * u < hi ? b0 + 4 * u : hi - 1;
            x[u] = part[(long long)b * ld + cc];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (b0 + 4 * u < hi) s += (double)x[u];
    }
    sh[sub][c] = col < ld ? s : 0.0;
    __syncthreads();
    if (sub == 0 && col < ld) pca_red(S, d)[(long long)g * ld + col] = sh[0][c] + sh[1][c] + sh[2][c] + sh[3][c];
}

// ---- PcaOrthonormalize: one block of 1024.  p = sum of the GR rows of red; classical
// Gram-Schmidt against the components before k (U = P_<k^T p, then p -= P_<k U: pca_cpu.cpp:356-372);
// p *= 1/|p| (dnrm2 + dscal: pca_cpu.cpp:374-376); P[k] = p.  EPT elements per thread (ld <=
// 1024 * EPT).  Loads read clamped, always valid addresses unconditionally, in groups of eight
// rows, so a group's loads issue together; each sum then takes its valid terms in order.
template <int EPT>
__device__ inline void pca_orthonormalize(float *__restrict__ S, const PCA_DIMS &d, double (&sW)[16][PTK_PCA_MAX_K],
                                          double (&sU)[PTK_PCA_MAX_K]) {
    constexpr int kGroup = 8;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, ld = d.ld, k = d.k, GR = d.GR;
    const double *red = pca_red(S, d);
    float *P = pca_P(S, d);
    int nc[EPT];
    bool nv[EPT];
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
        const int n = tid + 1024 * i;
        nv[i] = n < ld;
        nc[i] = n < ld ? n : ld - 1;
    }
    double p[EPT];
#pragma unroll
    for (int i = 0; i < EPT; ++i) p[i] = 0;
    for (int g0 = 0; g0 < GR; g0 += kGroup) {
        double rv[kGroup][EPT];
#pragma unroll
        for (int gg = 0; gg < kGroup; ++gg) {
            const int g = g0 + gg < GR ? g0 + gg : GR - 1;
#pragma unroll
            for (int i = 0; i < EPT; ++i) rv[gg][i] = red[(long long)g * ld + nc[i]];
        }
#pragma unroll
        for (int gg = 0; gg < kGroup; ++gg)
            if (g0 + gg < GR)
#pragma unroll
                for (int i = 0; i < EPT; ++i)
                    if (nv[i]) p[i] += rv[gg][i];
    }
    auto loadP = [&](float (&pv)[kGroup][EPT], int j0) {
#pragma unroll
        for (int jj = 0; jj < kGroup; ++jj) {
            const int j = j0 + jj < k ? j0 + jj : k - 1;
#pragma unroll
            for (int i = 0; i < EPT; ++i) pv[jj][i] = P[(long long)j * ld + nc[i]];
        }
    };
    if (k > 0) {
        // the k dot products eight at a time; the eight wave reductions run together
        for (int j0 = 0; j0 < k; j0 += kGroup) {
            float pv[kGroup][EPT];
            loadP(pv, j0);
            double v[kGroup];
#pragma unroll
            for (int jj = 0; jj < kGroup; ++jj) {
                v[jj] = 0;
                if (j0 + jj < k)
#pragma unroll
                    for (int i = 0; i < EPT; ++i)
                        if (nv[i]) v[jj] += (double)pv[jj][i] * p[i];
            }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
                for (int jj = 0; jj < kGroup; ++jj) v[jj] += __shfl_xor(v[jj], off, 64);
            if (lane == 0)
#pragma unroll
                for (int jj = 0; jj < kGroup; ++jj)
                    if (j0 + jj < k) sW[wv][j0 + jj] = v[jj];
        }
        __syncthreads();
        if (tid < k) {
            double u = 0;
            for (int w = 0; w < 16; ++w) u += sW[w][tid];
            sU[tid] = u;
        }
        __syncthreads();
        // p -= P_<k U: each element's sum over j in order
        double s[EPT];
#pragma unroll
        for (int i = 0; i < EPT; ++i) s[i] = 0;
        for (int j0 = 0; j0 < k; j0 += kGroup) {
            float pv[kGroup][EPT];
            loadP(pv, j0);
#pragma unroll
            for (int jj = 0; jj < kGroup; ++jj)
                if (j0 + jj < k)
#pragma unroll
                    for (int i = 0; i < EPT; ++i) s[i] += (double)pv[jj][i] * sU[j0 + jj];
        }
#pragma unroll
        for (int i = 0; i < EPT; ++i)
            if (nv[i]) p[i] -= s[i];
        __syncthreads();
    }
    double v = 0;  // sum of squares in element order (the terms beyond ld are zero)
#pragma unroll
    for (int i = 0; i < EPT; ++i) v = i == 0 ? p[0] * p[0] : v + p[i] * p[i];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) sW[wv][0] = v;
    __syncthreads();
    double nrm2 = 0;
    for (int w = 0; w < 16; ++w) nrm2 += sW[w][0];
    const double inv = 1.0 / sqrt(nrm2);
#pragma unroll
    for (int i = 0; i < EPT; ++i)
        if (nv[i]) P[(long long)k * ld + tid + 1024 * i] = (float)(p[i] * inv);
}

extern "C" __global__ void __launch_bounds__(1024)
PcaOrthonormalize(float *__restrict__ S, PCA_DIMS d) {
    __shared__ double sW[16][PTK_PCA_MAX_K];
    __shared__ double sU[PTK_PCA_MAX_K];
    if (!pca_ok(d) || blockDim.x != 1024 || gridDim.x != 1) return;
    if (d.ld <= 2048) pca_orthonormalize<2>(S, d, sW, sU);
    else pca_orthonormalize<4>(S, d, sW, sU);
}

// ---- PcaFinish: L_k = sqrt(sum of the G tn partials) (pca_cpu.cpp:385), one block of 256
// summing in a fixed order.
extern "C" __global__ void __launch_bounds__(256)
PcaFinish(float *__restrict__ S, PCA_DIMS d) {
    __shared__ double sh[256];
    if (!pca_ok(d) || blockDim.x != 256 || gridDim.x != 1) return;
    const int tid = threadIdx.x;
    const double *tn = pca_tn(S, d);
    double s = 0;
    for (int b = tid; b < d.G; b += 256) s += tn[b];
    sh[tid] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) sh[tid] += sh[tid + w];
        __syncthreads();
    }
    if (tid == 0) pca_L(S, d)[d.k] = sqrt(sh[0]);
}
#include "pca_blas.h"
