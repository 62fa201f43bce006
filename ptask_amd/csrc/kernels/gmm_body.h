// gmm_body.h -- EM for a diagonal Gaussian mixture: the block-level bodies of the two task
// kernels, shared by gmm.hip (one graph) and batch.hip (many instances).
//
//   GmmAccumulate(const float* X, GMM_DIMS d, const double* model, double* partials)
//     One pass over the points per EM step: the E-step (log responsibilities, per-point
//     log-sum-exp, log-likelihood) and the M-step's sufficient statistics, written as one row of
//     binary64 partial sums per block (no float atomics).
//   GmmFinalize(double* model, const double* partials, GMM_DIMS d)
//     Sums the rows in a fixed order (bitwise reproducible, whatever else shares the device),
//     applies the M-step and rewrites the binary32 E-step table of the model.
//
// Recurrence (ptaskUnitTest/nr3/gaumixmod.h:28-73, diagonal covariance):
//   r_nk = -1/2 sum_m (x_nm - mu_km)^2 / var_km - 1/2 sum_m log var_km + log frac_k,
//   tmp_n = max_k r + log sum_k exp(r_nk - max), resp = exp(r_nk - tmp_n), loglike = sum tmp_n,
//   w_k = sum_n resp, mu_k = sum resp x / w_k, var_k = sum resp (x - mu_k)^2 / w_k, frac = w/N.
// The moments are accumulated about the table's binary32 means c (the previous means):
//   su = sum r (x - c), suu = sum r (x - c)^2  =>  mu = c + su/w, var = suu/w - (su/w)^2,
// which is exact algebra and keeps the accumulated magnitudes near the spread of the component.
//
// GmmAccumulate, per 128-point tile (block of 256 = 4 waves):
//   E-phase: a lane owns one point (its coordinates in registers); waves 0-1 take components
//     [0, ceil(K/2)), waves 2-3 the rest, so the component index is uniform across a wave and the
//     table entries arrive through scalar loads, one VALU op per multiply-add and no LDS traffic.
//     Log responsibilities go to LDS; the two halves combine their maxima and sums through LDS.
//   M-phase: thread t owns (component, 4 dimensions) jobs t and t + 256 and walks the tile's
//     points, reading the responsibility and a float4 of coordinates from LDS.
// Tile sums are binary32 (128 points); the per-block running sums are binary64.
#pragma once

#include <hip/hip_runtime.h>

#include "ptk_gmm.h"


#define TILE 128
#define KMAX PTK_GMM_MAX_K
#define DMAX PTK_GMM_MAX_D
#define XS (DMAX + 4)    // LDS row stride of a point's coordinates, floats (16-byte rows)
#define RS (KMAX + 1)    // LDS row stride of a point's responsibilities, floats

__device__ inline bool gmm_dims_ok(const GMM_DIMS &d) {
    return d.N >= 1 && d.D >= 1 && d.D <= DMAX && d.K >= 1 && d.K <= KMAX && d.blocks >= 1 &&
           d.maxIter >= 0;
}

struct GmmShared {
    float sX[TILE * XS];
    float sR[TILE * RS];
    float sMax[2][TILE];
    float sSum[2][TILE];
    double sLL[4];
};

// Coordinates of tile t (zero for padding points and dimensions, nothing for t >= t1) into this
// thread's 2 NC registers: element i = tid + 256 j of the TILE x Dp tile.
template <int NC>
__device__ __forceinline__ void gmm_fetch_tile(const float *__restrict__ X, const GMM_DIMS &d, long long t,
                                               long long t1, float *pf) {
    const int Dp = 4 * NC, D = d.D;
    const long long p0 = t * TILE;
    const int np = t < t1 ? (int)min((long long)TILE, (long long)d.N - p0) : 0;
    const float *__restrict__ xt = X + p0 * D;
#pragma unroll
    for (int j = 0; j < 2 * NC; ++j) {
        const int i = (int)threadIdx.x + 256 * j, p = i / Dp, m = i - p * Dp;
        pf[j] = (p < np && m < D) ? xt[(long long)p * D + m] : 0.f;
    }
}

// NC = float4 chunks per point (Dp / 4), a compile-time constant so that the chunk loops unroll
// without guards and a component's table row arrives in one batch of scalar loads.
// Tiles t0, t0 + tstep, ... (< t1) of the points into one row of partial sums.
template <int NC>
__device__ __forceinline__ void gmm_accumulate_tiles(const float *__restrict__ X, const GMM_DIMS &d,
                                                     const double *__restrict__ model,
                                                     double *__restrict__ row, long long t0,
                                                     long long t1, long long tstep, GmmShared &sh) {
// Tolerance-checked binary32 work: let the compiler fuse multiply-adds (the build default is off).
#pragma clang fp contract(fast)
    float *sX = sh.sX, *sR = sh.sR;
    const int D = d.D, K = d.K, Dp = 4 * NC, nc = NC;
    const int tid = threadIdx.x, lane = tid & 63;
    const float *__restrict__ tab = (const float *)(model + ptk_gmm_table(&d));
    const float *__restrict__ Cf = tab + 2 * K * Dp;

    // E-phase ownership (wave-uniform component range)
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int ep = (wv & 1) * 64 + lane;
    const int eh = wv >> 1;
    const int kh = (K + 1) >> 1;
    const int k0 = eh * kh, k1 = min(K, k0 + kh);

    // M-phase ownership: jobs (component, chunk).  With more than 128 jobs thread t owns jobs t
    // and t + 256; with fewer, reps = 256 / jobs groups of threads each own every job and split
    // the tile's points (group g takes points g, g + reps, ...), their sums combined in a fixed
    // order at the end of the unit.
    const int jobs = K * nc;
    const int reps = jobs <= 128 ? 256 / jobs : 1;
    const int grp = reps > 1 ? tid / jobs : 0;
    int mk[2], mc[2];
    bool mv[2];
    float4 cen[2];
    double aw[2], asu[2][4], asuu[2][4];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int j = reps > 1 ? (q == 0 && tid < reps * jobs ? tid - grp * jobs : jobs) : tid + 256 * q;
        mv[q] = j < jobs;
        mk[q] = mv[q] ? j / nc : 0;
        mc[q] = mv[q] ? j - mk[q] * nc : 0;
        cen[q] = *(const float4 *)(tab + 2 * mk[q] * Dp + 4 * mc[q]);
        aw[q] = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) asu[q][e] = asuu[q][e] = 0;
    }
    double ll = 0;

    // The next tile's coordinates are loaded into registers while the current tile computes
    // (2 NC values per thread: TILE x Dp / 256), so the global-memory latency stays hidden.
    float pf[2 * NC];
    gmm_fetch_tile<NC>(X, d, t0, t1, pf);
    for (long long t = t0; t < t1; t += tstep) {
        const long long p0 = t * TILE;
        const int np = (int)min((long long)TILE, (long long)d.N - p0);
#pragma unroll
        for (int j = 0; j < 2 * NC; ++j) {
            const int i = tid + 256 * j, p = i / Dp, m = i - p * Dp;
            sX[p * XS + m] = pf[j];
        }
        __syncthreads();
        gmm_fetch_tile<NC>(X, d, t + tstep, t1, pf);

        // ---- E-phase: log responsibilities of point ep for components [k0, k1)
        float x[4 * NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const float4 v = *(const float4 *)&sX[ep * XS + 4 * c];
            x[4 * c] = v.x;
            x[4 * c + 1] = v.y;
            x[4 * c + 2] = v.z;
            x[4 * c + 3] = v.w;
        }
        float mx = -INFINITY;
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
            const float lp = fmaf(-0.5f, s0 + s1, Cf[k]);
            sR[ep * RS + k] = lp;
            mx = fmaxf(mx, lp);
        }
        sh.sMax[eh][ep] = mx;
        __syncthreads();
        mx = fmaxf(sh.sMax[0][ep], sh.sMax[1][ep]);
        float sum = 0.f;
        for (int k = k0; k < k1; ++k) sum += __expf(sR[ep * RS + k] - mx);
        sh.sSum[eh][ep] = sum;
        __syncthreads();
        const float tmp = mx + __logf(sh.sSum[0][ep] + sh.sSum[1][ep]);
        const bool valid = ep < np;
        for (int k = k0; k < k1; ++k) sR[ep * RS + k] = valid ? __expf(sR[ep * RS + k] - tmp) : 0.f;
        if (valid && eh == 0) ll += (double)tmp;
        __syncthreads();

        // ---- M-phase: tile sums for jobs (mk, mc)
        float w[2], su[2][4], suu[2][4];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            w[q] = 0.f;
#pragma unroll
            for (int e = 0; e < 4; ++e) su[q][e] = suu[q][e] = 0.f;
        }
        auto point = [&](int p) {
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const float r = sR[p * RS + mk[q]];
                const float4 xv = *(const float4 *)&sX[p * XS + 4 * mc[q]];
                const float u[4] = {xv.x - cen[q].x, xv.y - cen[q].y, xv.z - cen[q].z, xv.w - cen[q].w};
                w[q] += r;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const float ru = r * u[e];
                    su[q][e] += ru;
                    suu[q][e] = fmaf(ru, u[e], suu[q][e]);
                }
            }
        };
        if (reps == 1) {  // the common large-K case keeps a unit-stride loop
            for (int p = 0; p < np; ++p) point(p);
        } else {
            for (int p = grp; p < np; p += reps) point(p);
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            aw[q] += (double)w[q];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                asu[q][e] += (double)su[q][e];
                asuu[q][e] += (double)suu[q][e];
            }
        }
        __syncthreads();
    }

    if (reps > 1) {  // combine the groups' sums in group order (LDS of sR is free now)
        double *acc = (double *)sR;
        if (mv[0] || tid < reps * jobs) {
            double *a = acc + (size_t)tid * 9;
            a[0] = aw[0];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                a[1 + e] = asu[0][e];
                a[5 + e] = asuu[0][e];
            }
        }
        __syncthreads();
        if (grp == 0 && mv[0]) {
#pragma unroll
            for (int e = 0; e < 9; ++e) {
                double v = 0;
                for (int g = 0; g < reps; ++g) v += acc[(size_t)(g * jobs + tid) * 9 + e];
                if (e == 0) aw[0] = v;
                else if (e < 5) asu[0][e - 1] = v;
                else asuu[0][e - 5] = v;
            }
        }
        if (grp != 0) mv[0] = false;
        __syncthreads();
    }
    // the row of partial sums
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        if (!mv[q]) continue;
        const int k = mk[q];
        if (mc[q] == 0) row[k] = aw[q];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int m = 4 * mc[q] + e;
            if (m < D) {
                row[K + k * D + m] = asu[q][e];
                row[K + K * D + k * D + m] = asuu[q][e];
            }
        }
    }
    double v = ll;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) sh.sLL[wv] = v;
    __syncthreads();
    if (tid == 0) row[K + 2 * K * D] = sh.sLL[0] + sh.sLL[1] + sh.sLL[2] + sh.sLL[3];
    __syncthreads();  // a block may go on to another unit, which reuses sLL
}

// Row b of the partial sums covers the tiles b, b + blocks, b + 2*blocks, ...: a fixed function
// of (N, blocks), whoever runs it, so every launch form (one graph, a batch of instances)
// produces bitwise the same partial sums; blocks running at the same time read neighbouring
// tiles.
__device__ __forceinline__ void gmm_accumulate_unit(const float *__restrict__ X, const GMM_DIMS &d,
                                                    const double *__restrict__ model,
                                                    double *__restrict__ partials, int b, GmmShared &sh) {
    const long long t0 = b, t1 = ((long long)d.N + TILE - 1) / TILE, ts = d.blocks;
    double *row = partials + (long long)b * ptk_gmm_partial_stride(&d);
    switch (ptk_gmm_dpad(&d) >> 2) {
        case 1: gmm_accumulate_tiles<1>(X, d, model, row, t0, t1, ts, sh); break;
        case 2: gmm_accumulate_tiles<2>(X, d, model, row, t0, t1, ts, sh); break;
        case 3: gmm_accumulate_tiles<3>(X, d, model, row, t0, t1, ts, sh); break;
        case 4: gmm_accumulate_tiles<4>(X, d, model, row, t0, t1, ts, sh); break;
        case 5: gmm_accumulate_tiles<5>(X, d, model, row, t0, t1, ts, sh); break;
        case 6: gmm_accumulate_tiles<6>(X, d, model, row, t0, t1, ts, sh); break;
        case 7: gmm_accumulate_tiles<7>(X, d, model, row, t0, t1, ts, sh); break;
        default: gmm_accumulate_tiles<8>(X, d, model, row, t0, t1, ts, sh); break;
    }
}

// Grid K + 1 blocks of 256.  Block k < K sums the 2D + 1 partial columns of component k over the
// rows (256 / (2D + 1) interleaved row groups, combined in a fixed order), applies the M-step and
// rewrites the component's table entries; block K sums the log-likelihood and advances the step
// counter.
__device__ inline void gmm_finalize_block(double *__restrict__ model, const double *__restrict__ partials,
                                          const GMM_DIMS &d, int k) {
    __shared__ double sTot[256];
    __shared__ double sLn[DMAX];
    __shared__ double sRed[256];
    const int D = d.D, K = d.K, P = d.blocks, Dp = ptk_gmm_dpad(&d);
    const long long S = ptk_gmm_partial_stride(&d);
    const int tid = threadIdx.x;
    if (k < K) {
        const int ncol = 2 * D + 1, G = 256 / ncol;
        const int g = tid / ncol, c = tid - g * ncol;
        double acc = 0;
        if (g < G) {
            const long long col = c == 0   ? (long long)k
                                  : c <= D ? K + (long long)k * D + (c - 1)
                                           : K + (long long)K * D + (long long)k * D + (c - 1 - D);
            // rows g, g + G, ... in order; sixteen rows' loads issued together (a row past the
            // last reads row g and is not added)
            const double *src = partials + col;
            for (int b0 = g; b0 < P; b0 += 16 * G) {
                double v[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const int b = b0 + u * G;
                    v[u] = src[(long long)(b < P ? b : g) * S];
                }
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    if (b0 + u * G < P) acc += v[u];
            }
        }
        sTot[tid] = acc;
        __syncthreads();
        double t = 0;  // the groups in order; column c's total then kept in sTot[c]
        if (tid < ncol)
            for (int q = 0; q < G; ++q) t += sTot[q * ncol + tid];
        __syncthreads();
        if (tid < ncol) sTot[tid] = t;
        __syncthreads();
        double *mu = model + ptk_gmm_mu(&d), *var = model + ptk_gmm_var(&d);
        float *tab = (float *)(model + ptk_gmm_table(&d));
        const double w = sTot[0];
        if (tid < D) {
            const int m = tid;
            const long long i = (long long)k * D + m;
            double mnew = mu[i], vnew = var[i];
            if (w != 0) {
                const double cen = (double)tab[2 * k * Dp + m];
                const double dm = sTot[1 + m] / w;
                mnew = cen + dm;
                vnew = sTot[1 + D + m] / w - dm * dm;
            }
            mu[i] = mnew;
            var[i] = vnew;
            tab[2 * k * Dp + m] = (float)mnew;
            tab[2 * k * Dp + Dp + m] = (float)(1.0 / vnew);
            sLn[m] = log(vnew);
        }
        __syncthreads();
        if (tid == 0) {
            double lndet = 0;
            for (int m = 0; m < D; ++m) lndet += sLn[m];
            const double f = w / (double)d.N;
            model[ptk_gmm_frac(&d) + k] = f;
            tab[2 * K * Dp + k] = (float)(-0.5 * lndet + log(f));
        }
    } else {
        double acc = 0;
        for (int b = tid; b < P; b += 256) acc += partials[(long long)b * S + K + 2ll * K * D];
        sRed[tid] = acc;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (tid < s) sRed[tid] += sRed[tid + s];
            __syncthreads();
        }
        if (tid == 0) {
            const long long step = (long long)model[0];
            if (step >= 0 && step <= d.maxIter) model[ptk_gmm_ll(&d) + step] = sRed[0];
            model[0] = (double)(step + 1);
        }
    }
    __syncthreads();
}
