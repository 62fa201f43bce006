// gmm_mfma.h -- GmmAccumulate with its dense parts on the matrix cores (v_mfma_f32_16x16x4f32,
// and v_mfma_f32_32x32x2f32 for the M-step when its moments fill four 32 x 32 tiles).
//
//   GmmAccumulateMfma(const float* X, GMM_DIMS d, const double* model, double* partials)
//     Same arguments, partial-row format and tile-to-row assignment as GmmAccumulate
//     (gmm_body.h), so GmmFinalize and every caller are unchanged; only the rounding differs.
//
// Per 128-point tile, with coordinates centred on g (the mean of the component means, one
// per-dimension vector per launch: x' = x - g, c'_k = mu_k - g):
//   E-step  the squared Mahalanobis distances of all points to all components as one product
//           [x'^2 | x'] (128 x 2Dp)  .  [iv ; -2 iv c'] (2Dp x Kp)  + sum_m iv c'^2,
//           then log responsibilities lp = C_k - d/2 and the per-point log-sum-exp on the VALU.
//           The expansion cancels (|x'|^2 iv against the distance); it runs only while every
//           reciprocal variance is at most GMM_MFMA_IV_LIMIT.  Otherwise (the first EM step,
//           var = 1e-10) the block keeps GmmAccumulate's per-component differences on the VALU.
//   M-step  the moments about g as one product  R^T (Kp x 128) . [x' | x'^2] (128 x 2Dp);
//           W_k = sum_p R_pk on the VALU.  Tile sums are binary32 (in MFMA order), added into
//           binary64 per block, and converted once per block to moments about the table centres
//           c = mu_k:  su = A1 - c' W,  suu = A2 - 2 c' A1 + c'^2 W  (binary64).
// Output tile ownership is fixed per wave, so each partial row is a fixed function of (N,
// blocks): bitwise reproducible run to run.
#pragma once

#include "gmm_body.h"

typedef float gmm_f4 __attribute__((ext_vector_type(4)));

#define GMM_MFMA_IV_LIMIT 1.0e3f
#define MXS (DMAX + 4)  // floats per point row of centred coordinates
#define MRS (KMAX + 5)  // floats per point row of (log) responsibilities: odd, so a wave reading
                        // one component of 64 points hits distinct banks
#define MBS (KMAX + 4)  // floats per row of the E-step B operand

struct GmmMfmaShared {
    union {
        struct {
            float sX[TILE * MXS];
            float sR[TILE * MRS];
        } t;
        double sAcc[KMAX * 2 * DMAX];  // the block's moment sums about g, after the last tile
    } u;
    float sB[2 * DMAX * MBS];
    float sE[KMAX];
    float sG[DMAX];
    float sMax[2][TILE];
    float sSum[2][TILE];
    float sWp[4][KMAX];  // per-wave column sums of a tile's responsibilities
    double sW[KMAX];
    double sLL[4];
};

// E-step products of component tile kt with point tiles IL pg .. IL pg + IL - 1 of the tile in
// sX: IL independent MFMA chains, interleaved, sharing each B operand load; log responsibilities
// C_k - d/2 into sR.
template <int IL, int NC>
__device__ __forceinline__ void gmm_mfma_estep(GmmMfmaShared &sh, const float *__restrict__ sX, float *__restrict__ sR,
                                               int kt, int pg, int K, int l16, int l4) {
#pragma clang fp contract(fast)
    constexpr int Dp = 4 * NC, D2 = 2 * Dp;
    const float *bb = sh.sB + kt * 16 + l16;
    gmm_f4 acc[IL];
#pragma unroll
    for (int u = 0; u < IL; ++u) acc[u] = gmm_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < D2 / 4; ++s) {
        const int c = 4 * s + l4, cx = c < Dp ? c : c - Dp;
        const float bv = bb[c * MBS];
#pragma unroll
        for (int u = 0; u < IL; ++u) {
            const float xv = sX[((IL * pg + u) * 16 + l16) * MXS + cx];
            acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(c < Dp ? xv * xv : xv, bv, acc[u], 0, 0, 0);
        }
    }
    const int k = kt * 16 + l16;
    if (k < K) {
        const float e = sh.sE[k];
#pragma unroll
        for (int u = 0; u < IL; ++u)
#pragma unroll
            for (int r = 0; r < 4; ++r) sR[((IL * pg + u) * 16 + 4 * l4 + r) * MRS + k] = fmaf(-0.5f, acc[u][r], e);
    }
}

template <int NC>
__device__ __forceinline__ void gmm_mfma_tiles(const float *__restrict__ X, const GMM_DIMS &d,
                                               const double *__restrict__ model, double *__restrict__ row,
                                               long long t0, long long t1, long long tstep, GmmMfmaShared &sh) {
#pragma clang fp contract(fast)
    float *sX = sh.u.t.sX, *sR = sh.u.t.sR;
    const int D = d.D, K = d.K, Dp = 4 * NC, D2 = 2 * Dp, Kp = (K + 15) & ~15;
    const int KT = Kp >> 4, JT = (D2 + 15) >> 4;
    // M-step on 32 x 32 x 2 products when the moments fill four 32 x 32 tiles (K and 2 Dp in
    // (32, 64]): one tile per wave, half the operand reads per flop of the 16 x 16 x 4 form
    const int JT2 = (D2 + 31) >> 5;
    const bool m32 = ((K + 31) >> 5) * JT2 == 4;
    typedef float gmm_f16 __attribute__((ext_vector_type(16)));
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int l16 = lane & 15, l4 = lane >> 4;
    const float *__restrict__ tab = (const float *)(model + ptk_gmm_table(&d));
    const float *__restrict__ Cf = tab + 2 * K * Dp;

    // ---- per launch: centre, E-step operand and constants
    for (int m = tid; m < Dp; m += 256) {
        double s = 0;
        for (int k = 0; k < K; ++k) s += (double)tab[2 * k * Dp + m];
        sh.sG[m] = m < D ? (float)(s / K) : 0.f;
    }
    __syncthreads();
    bool big = false;
    for (int i = tid; i < D2 * Kp; i += 256) {
        const int c = i / Kp, k = i - c * Kp;
        float v = 0.f;
        if (k < K) {
            const int m = c < Dp ? c : c - Dp;
            const float iv = tab[2 * k * Dp + Dp + m];
            if (m < D && !(iv <= GMM_MFMA_IV_LIMIT)) big = true;
            v = c < Dp ? iv : -2.f * iv * (tab[2 * k * Dp + m] - sh.sG[m]);
        }
        sh.sB[c * MBS + k] = v;
    }
    for (int k = tid; k < K; k += 256) {
        double q = 0;
        for (int m = 0; m < D; ++m) {
            const double cp = (double)tab[2 * k * Dp + m] - (double)sh.sG[m];
            q += (double)tab[2 * k * Dp + Dp + m] * cp * cp;
        }
        sh.sE[k] = (float)((double)Cf[k] - 0.5 * q);
    }
    const bool mfmaE = !__syncthreads_or(big);

    // E-phase ownership of the VALU fallback and of the log-sum-exp: point ep, components of half eh
    const int ep = (wv & 1) * 64 + lane, eh = wv >> 1;
    const int kh = (K + 1) >> 1, k0 = eh * kh, k1 = min(K, k0 + kh);

    double dacc[4][4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int r = 0; r < 4; ++r) dacc[q][r] = 0;
    double aw = 0, ll = 0;

    // next tile's coordinates in registers while the current one computes (gmm_fetch_tile)
    float pf[2 * NC];
    gmm_fetch_tile<NC>(X, d, t0, t1, pf);
    for (long long t = t0; t < t1; t += tstep) {
        const long long p0 = t * TILE;
        const int np = (int)min((long long)TILE, (long long)d.N - p0);
        const float *__restrict__ xt = X + p0 * D;
#pragma unroll
        for (int j = 0; j < 2 * NC; ++j) {
            const int i = tid + 256 * j, p = i / Dp, m = i - p * Dp;
            sX[p * MXS + m] = (p < np && m < D) ? pf[j] - sh.sG[m] : 0.f;
        }
        __syncthreads();
        gmm_fetch_tile<NC>(X, d, t + tstep, t1, pf);

        // ---- E-step: log responsibilities into sR[p][k]
        if (mfmaE) {
            // job jb: component tile jb % KT against point tiles 4 (jb / KT) .. +3, four
            // independent MFMA chains interleaved (enough to cover the product's latency)
            for (int jb = wv; jb < 2 * KT; jb += 4) gmm_mfma_estep<4, NC>(sh, sX, sR, jb % KT, jb / KT, K, l16, l4);
        } else {
            // differences to each component mean from the point itself, as GmmAccumulate
            float x[4 * NC];
#pragma unroll
            for (int m = 0; m < 4 * NC; ++m) x[m] = (ep < np && m < D) ? xt[(long long)ep * D + m] : 0.f;
            for (int k = k0; k < k1; ++k) {
                const float *__restrict__ tk = tab + 2 * k * Dp;
                float s0 = 0.f, s1 = 0.f;
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    const float4 mu = *(const float4 *)(tk + 4 * c);
                    const float4 iv = *(const float4 *)(tk + Dp + 4 * c);
                    const float u0 = x[4 * c] - mu.x, u1 = x[4 * c + 1] - mu.y;
                    const float u2 = x[4 * c + 2] - mu.z, u3 = x[4 * c + 3] - mu.w;
                    s0 = fmaf(u0 * u0, iv.x, s0);
                    s1 = fmaf(u1 * u1, iv.y, s1);
                    s0 = fmaf(u2 * u2, iv.z, s0);
                    s1 = fmaf(u3 * u3, iv.w, s1);
                }
                sR[ep * MRS + k] = fmaf(-0.5f, s0 + s1, Cf[k]);
            }
        }
        __syncthreads();

        // ---- per-point log-sum-exp, responsibilities, log-likelihood
        float mx = -INFINITY;
#pragma unroll 8
        for (int k = k0; k < k1; ++k) mx = fmaxf(mx, sR[ep * MRS + k]);
        sh.sMax[eh][ep] = mx;
        __syncthreads();
        mx = fmaxf(sh.sMax[0][ep], sh.sMax[1][ep]);
        float sum = 0.f;
#pragma unroll 8
        for (int k = k0; k < k1; ++k) sum += __expf(sR[ep * MRS + k] - mx);
        sh.sSum[eh][ep] = sum;
        __syncthreads();
        const float tmp = mx + __logf(sh.sSum[0][ep] + sh.sSum[1][ep]);
        const bool valid = ep < np;
#pragma unroll 8
        for (int k = k0; k < k1; ++k) sR[ep * MRS + k] = valid ? __expf(sR[ep * MRS + k] - tmp) : 0.f;
        if (valid && eh == 0) ll += (double)tmp;
        __syncthreads();

        // ---- M-step: R^T [x' | x'^2] on the matrix cores, W on the VALU.
        if (m32) {
            // wave wv: components 32 (wv / JT2) .., features 32 (wv % JT2) ..; lane l supplies
            // A[l % 32][l / 32] = R[p][k] and B[l / 32][l % 32] = [x' | x'^2][p][j] for the two
            // points p = 2s + l / 32 of step s; D[8(r/4) + 4(l/32) + r%4][l % 32] in acc[r]
            const int kt2 = wv / JT2, jt2 = wv - kt2 * JT2, l32 = lane & 31, h = lane >> 5;
            const int ki = kt2 * 32 + l32, j = jt2 * 32 + l32;
            const bool rowok = ki < K, jlo = j < Dp, jhi = j >= Dp && j < D2;
            const int jx = j < Dp ? j : (j < D2 ? j - Dp : 0);
            gmm_f16 acc = {};
#pragma unroll 8
            for (int s = 0; s < TILE / 2; ++s) {
                const int p = 2 * s + h;
                const float a = rowok ? sR[p * MRS + ki] : 0.f;
                const float xv = sX[p * MXS + jx];
                const float b = jlo ? xv : (jhi ? xv * xv : 0.f);
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int r = 0; r < 4; ++r) dacc[q][r] += (double)acc[4 * q + r];
        } else {
            // the wave's four 16 x 16 output tiles (tt = wv + 4q) accumulated together, their
            // products interleaved; a tile past the last (tt >= KT JT) multiplies zeros and is
            // not stored
            int ki[4], jx[4];
            bool rowok[4], jlo[4], jhi[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int tt = wv + 4 * q;
                const bool on = tt < KT * JT;
                const int kt = on ? tt / JT : 0, jt = on ? tt - kt * JT : 0;
                const int j = jt * 16 + l16;
                ki[q] = kt * 16 + l16;
                rowok[q] = on && ki[q] < K;
                jlo[q] = on && j < Dp;
                jhi[q] = on && j >= Dp && j < D2;
                jx[q] = j < Dp ? j : (j < D2 ? j - Dp : 0);
            }
            gmm_f4 acc[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] = gmm_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
            for (int s = 0; s < TILE / 4; ++s) {
                const int p = 4 * s + l4;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float a = rowok[q] ? sR[p * MRS + ki[q]] : 0.f;
                    const float xv = sX[p * MXS + jx[q]];
                    const float b = jlo[q] ? xv : (jhi[q] ? xv * xv : 0.f);
                    acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[q], 0, 0, 0);
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (wv + 4 * q < KT * JT)
#pragma unroll
                    for (int r = 0; r < 4; ++r) dacc[q][r] += (double)acc[q][r];
        }
        {   // W: wave wv sums points [32 wv, 32 wv + 32) of component `lane`; the four wave sums
            // are added in wave order below, after the barrier
            float w = 0.f;
            if (lane < K) {
#pragma unroll 8
                for (int p = 32 * wv; p < 32 * wv + 32; ++p) w += sR[p * MRS + lane];
            }
            sh.sWp[wv][lane] = w;
        }
        __syncthreads();
        if (tid < K) aw += (double)(((sh.sWp[0][tid] + sh.sWp[1][tid]) + sh.sWp[2][tid]) + sh.sWp[3][tid]);
    }

    // ---- the block's moments about the table centres, one partial row
    double *sAcc = sh.u.sAcc;
    if (m32) {
        const int kt2 = wv / JT2, jt2 = wv - kt2 * JT2, j = jt2 * 32 + (lane & 31);
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = kt2 * 32 + 8 * q + 4 * (lane >> 5) + r;
                if (k < K && j < D2) sAcc[k * D2 + j] = dacc[q][r];
            }
    } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int tt = wv + 4 * q;
            if (tt < KT * JT) {
                const int kt = tt / JT, jt = tt - kt * JT, j = jt * 16 + l16;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int k = kt * 16 + 4 * l4 + r;
                    if (k < K && j < D2) sAcc[k * D2 + j] = dacc[q][r];
                }
            }
        }
    }
    if (tid < K) sh.sW[tid] = aw;
    __syncthreads();
    for (int i = tid; i < K * D; i += 256) {
        const int k = i / D, m = i - k * D;
        const double w = sh.sW[k], a1 = sAcc[k * D2 + m], a2 = sAcc[k * D2 + Dp + m];
        const double cp = (double)tab[2 * k * Dp + m] - (double)sh.sG[m];
        row[K + k * D + m] = a1 - cp * w;
        row[K + K * D + k * D + m] = a2 - 2.0 * cp * a1 + cp * cp * w;
    }
    if (tid < K) row[tid] = aw;
    double v = ll;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) sh.sLL[wv] = v;
    __syncthreads();
    if (tid == 0) row[K + 2 * K * D] = sh.sLL[0] + sh.sLL[1] + sh.sLL[2] + sh.sLL[3];
    __syncthreads();
}

__device__ __forceinline__ void gmm_mfma_unit(const float *__restrict__ X, const GMM_DIMS &d,
                                              const double *__restrict__ model, double *__restrict__ partials,
                                              int b, GmmMfmaShared &sh) {
    const long long t0 = b, t1 = ((long long)d.N + TILE - 1) / TILE, ts = d.blocks;
    double *row = partials + (long long)b * ptk_gmm_partial_stride(&d);
    switch (ptk_gmm_dpad(&d) >> 2) {
        case 1: gmm_mfma_tiles<1>(X, d, model, row, t0, t1, ts, sh); break;
        case 2: gmm_mfma_tiles<2>(X, d, model, row, t0, t1, ts, sh); break;
        case 3: gmm_mfma_tiles<3>(X, d, model, row, t0, t1, ts, sh); break;
        case 4: gmm_mfma_tiles<4>(X, d, model, row, t0, t1, ts, sh); break;
        case 5: gmm_mfma_tiles<5>(X, d, model, row, t0, t1, ts, sh); break;
        case 6: gmm_mfma_tiles<6>(X, d, model, row, t0, t1, ts, sh); break;
        case 7: gmm_mfma_tiles<7>(X, d, model, row, t0, t1, ts, sh); break;
        default: gmm_mfma_tiles<8>(X, d, model, row, t0, t1, ts, sh); break;
    }
}
