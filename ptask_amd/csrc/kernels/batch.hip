// batch.hip -- persistent multi-instance task kernels of the batched / served executor
// (code object batch.hsaco; the executor is ptask_amd/csrc/batch/executor.cpp).
//
// Every kernel takes a work list in device memory, the list's length in device memory and the
// array of instance descriptors (include/ptk_batch.h), and strides over the list with whatever
// grid it was launched with.  Because lengths live on the device, one captured hipGraph of these
// launches stays valid while instances are admitted and retired: the host only rewrites the
// lists between iterations.  The per-cell / per-row / per-tile arithmetic is the single-graph
// kernels' own (fdtd_halves.h, groupby_body.h, gmm_body.h, gmm_mfma.h), so an instance computes bitwise what
// its graph computes alone.
//
//   BatchFdtdH / BatchFdtdE   tiles {inst, col0, ncols, k0 | k1 << 16}: H then E half-step of
//                             the columns [col0, col0+ncols) and planes [k0, k1) of an instance;
//                             set (done & 1) is current, the other set receives the new fields
//   BatchFdtdAdvance          active FDTD instances: done += 1
//   BatchFdtdStep             tiles {inst, tile x, tile y, k0 | k1 << 16}: a whole step on
//                             overlapped 62 x 6-column tiles (the form when fdtd_cols == 0)
//   BatchFdtdSteps3           items {inst, tile x, tile y, k0 | k1 << 16} over 58 x 10 columns in
//                             per-block ranges: three whole steps per launch
//   BatchGroupbyHistogram/Partition           units {inst, unit}: chunkRows rows
//   BatchGroupbyBucketScan                    units {inst, key bucket}
//   BatchGroupbyKeyCount/Reshuffle            units {inst, tile}
//   BatchGroupbyKeyScan                       units {inst, half bucket}
//   BatchGroupbyTotals                        units {inst, quarter bucket}
//   BatchGmmAccumulate        units {inst, partial row b}
//   BatchGmmAccumulateMfma    the same units on the matrix cores (gmm_mfma.h; the executor's)
//   BatchGmmFinalize          units {inst, component k (K = log-likelihood block)}
//   BatchInit*                device-side input generation at admission
//   BatchChecksum*            64-bit checksums of results (sums of mixed words: the order in
//                             which they are added does not matter)
#include <hip/hip_runtime.h>

#include "fdtd_common.h"
#include "fdtd_halves.h"
#include "gmm_body.h"
#include "gmm_mfma.h"
// the work-list loop costs the keyscan a few registers: 32 tiles per load batch keep it unspilled
#define GB_KS_BATCH 32
#include "groupby_body.h"
#include "ptk_batch.h"
#include "ptk_prng.h"

using namespace ptk_fdtd;
using namespace ptk_groupby;

// ---- FDTD -------------------------------------------------------------------------------------

extern "C" __global__ void __launch_bounds__(256)
BatchFdtdH(const int4 *__restrict__ tiles, const int *__restrict__ count, const PTK_FDTD_INST *__restrict__ insts) {
    const int n = *count;
    for (int t = blockIdx.x; t < n; t += gridDim.x) {
        const int4 tl = tiles[t];
        const PTK_FDTD_INST &I = insts[tl.x];
        const int cur = I.done & 1, nxt = cur ^ 1;
        const HalfH F{I.f[cur][0], I.f[cur][1], I.f[cur][2], I.f[cur][3], I.f[cur][4], I.f[cur][5],
                      I.f[nxt][0], I.f[nxt][1], I.f[nxt][2]};
        const Ix X{I.p.Nx, I.p.Ny, I.p.Nz};
        const Coef c = make_coef(I.p);
        const int k0 = tl.w & 0xffff, k1 = (int)((unsigned)tl.w >> 16);
        for (int q = tl.y + (int)threadIdx.x; q < tl.y + tl.z; q += blockDim.x) h_column(F, X, c, (unsigned)q, k0, k1);
    }
}

extern "C" __global__ void __launch_bounds__(256)
BatchFdtdE(const int4 *__restrict__ tiles, const int *__restrict__ count, const PTK_FDTD_INST *__restrict__ insts) {
    const int n = *count;
    for (int t = blockIdx.x; t < n; t += gridDim.x) {
        const int4 tl = tiles[t];
        const PTK_FDTD_INST &I = insts[tl.x];
        const int cur = I.done & 1, nxt = cur ^ 1;
        const HalfE F{I.f[cur][3], I.f[cur][4], I.f[cur][5], I.f[nxt][0], I.f[nxt][1], I.f[nxt][2],
                      I.f[nxt][3], I.f[nxt][4], I.f[nxt][5]};
        const Ix X{I.p.Nx, I.p.Ny, I.p.Nz};
        const Coef c = make_coef(I.p);
        const int k0 = tl.w & 0xffff, k1 = (int)((unsigned)tl.w >> 16);
        for (int q = tl.y + (int)threadIdx.x; q < tl.y + tl.z; q += blockDim.x) e_column(F, X, c, (unsigned)q, k0, k1);
    }
}

// One whole Yee step on overlapped tiles (overlap_tile, fdtd_fused.h): tiles {inst, tile x, tile
// y, k0 | k1 << 16} of 62 x 6 owned columns; blocks of 64 x 8 threads.  Tiles are taken in an
// XCD-contiguous order so neighbouring tiles share their XCD's L2.  Instances must satisfy
// fits_buffer (the executor checks before it builds these tiles).
extern "C" __global__ void __launch_bounds__(512)
BatchFdtdStep(const int4 *__restrict__ tiles, const int *__restrict__ count, const PTK_FDTD_INST *__restrict__ insts) {
    __shared__ OverlapLds<8> L;
    if (blockDim.x != 64 || blockDim.y != 8 || blockDim.z != 1) return;  // geometry contract
    // a persistent grid larger than the list: the first n blocks take one tile each, spread
    // evenly over the XCDs (the extra blocks return)
    const int n = *count, G = (int)gridDim.x < n ? (int)gridDim.x : n;
    if ((int)blockIdx.x >= G) return;
    for (int t = (int)gb_xcd_block(blockIdx.x, (unsigned)G); t < n; t += G) {
        const int4 tl = tiles[t];
        const PTK_FDTD_INST &I = insts[tl.x];
        if (!fits_buffer(I.p)) continue;
        const int cur = I.done & 1, nxt = cur ^ 1;
        const Fields F{I.f[cur][0], I.f[cur][1], I.f[cur][2], I.f[cur][3], I.f[cur][4], I.f[cur][5],
                       I.f[nxt][0], I.f[nxt][1], I.f[nxt][2], I.f[nxt][3], I.f[nxt][4], I.f[nxt][5]};
        const int k0 = tl.w & 0xffff, k1 = (int)((unsigned)tl.w >> 16);
        overlap_tile<8>(F, I.p, tl.y * 62, tl.z * 6, k0, k1, L);
        lds_barrier();  // the next tile reuses LDS
    }
}

// Three whole Yee steps per launch on overlapped tiles (overlapn_tile<3>, fdtd_fused.h) over a
// persistent schedule of G virtual blocks (count[1]): virtual block b's items are
// [starts[b], starts[b+1]) of `tiles`, each {inst, tile x, tile y, k0 | k1 << 16} of 58 x 10
// owned columns, taken in order.  The executor builds the schedule (whole-depth tiles round-robin,
// then equal runs of the remaining tiles' planes); blocks of 64 x 16 threads take virtual blocks
// in grid strides, XCD-contiguously, so any grid size covers every item once.  A launch reads
// set (done & 1) and writes the other set three steps later, so one BatchFdtdAdvance per launch
// keeps the set index right.  Same instance conditions as BatchFdtdStep.
extern "C" __global__ void __launch_bounds__(1024)
BatchFdtdSteps3(const int4 *__restrict__ tiles, const int *__restrict__ count, const int *__restrict__ starts,
                const PTK_FDTD_INST *__restrict__ insts) {
    __shared__ OverlapNLds<3, 16, 1> L;
    if (blockDim.x != 64 || blockDim.y != 16 || blockDim.z != 1) return;  // geometry contract
    const int G = count[1];
    for (int vb = (int)blockIdx.x; vb < G; vb += (int)gridDim.x) {
        const int b = (int)gb_xcd_block((unsigned)vb, (unsigned)G);
        for (int t = starts[b]; t < starts[b + 1]; ++t) {
            const int4 tl = tiles[t];
            const PTK_FDTD_INST &I = insts[tl.x];
            if (!fits_buffer(I.p) || I.p.Nx < 1 || I.p.Ny < 1 || I.p.Nz < 1) continue;
            const int cur = I.done & 1, nxt = cur ^ 1;
            const Fields F{I.f[cur][0], I.f[cur][1], I.f[cur][2], I.f[cur][3], I.f[cur][4], I.f[cur][5],
                           I.f[nxt][0], I.f[nxt][1], I.f[nxt][2], I.f[nxt][3], I.f[nxt][4], I.f[nxt][5]};
            const int k0 = tl.w & 0xffff, k1 = (int)((unsigned)tl.w >> 16);
            overlapn_tile<3, 16>(F, I.p, tl.y * 58, tl.z * 10, k0, k1, L);
            lds_barrier();  // the next item reuses LDS
        }
    }
}

extern "C" __global__ void __launch_bounds__(256)
BatchFdtdAdvance(const int *__restrict__ list, const int *__restrict__ count, PTK_FDTD_INST *insts) {
    const int n = *count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) insts[list[i]].done += 1;
}

// The six fields of set A of one instance, U[-1,1) as FdtdInitField draws them (stream = field).
extern "C" __global__ void __launch_bounds__(256)
BatchInitFdtd(const PTK_FDTD_INST *__restrict__ insts, int inst, unsigned long long seed) {
    const PTK_FDTD_INST &I = insts[inst];
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (int f = 0; f < 6; ++f) {
        int d1, d2, d3;
        ptk_fdtd_extent(f, I.p.Nx, I.p.Ny, I.p.Nz, &d1, &d2, &d3);
        const unsigned long long n = (unsigned long long)d1 * d2 * d3;
        float *a = I.f[0][f];
        for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
            a[i] = ptk_rand_sym_f32(seed, (uint64_t)f, i);
    }
}

// ---- groupby ----------------------------------------------------------------------------------
// The groupby graph's seven passes over work lists; each pass's units are independent, and the
// passes run in order (the executor launches them in sequence).

// Units {inst, u}: the chunks of rows [u * chunkRows, (u + 1) * chunkRows).
extern "C" __global__ void __launch_bounds__(PTK_GB_HIST_THREADS)
BatchGroupbyHistogram(const int2 *__restrict__ units, const int *__restrict__ count, const PTK_GROUPBY_INST *__restrict__ insts) {
    __shared__ GbHistLds L;
    if (blockDim.x != PTK_GB_HIST_THREADS) return;
    const int n = *count;
    for (int u = blockIdx.x; u < n; u += gridDim.x) {
        const int2 w = units[u];
        const PTK_GROUPBY_INST &I = insts[w.x];
        if (!gb_keys_ok(I.numKeys) || I.n <= 0) continue;
        const PTK_GB_SCRATCH S = ptk_gb_scratch(I.scratch, I.n, I.numKeys);
        const long long per = I.chunkRows / PTK_GB_CHUNK, c0 = (long long)w.y * per;
        for (long long c = c0; c < c0 + per && c < S.chunks; ++c) gb_histogram_chunk(I.keys, I.n, I.numKeys, c, S, L);
    }
}

// Units {inst, bucket}.
extern "C" __global__ void __launch_bounds__(PTK_GB_SCAN_THREADS)
BatchGroupbyBucketScan(const int2 *__restrict__ units, const int *__restrict__ count, const PTK_GROUPBY_INST *__restrict__ insts) {
    __shared__ unsigned red[16];
    if (blockDim.x != PTK_GB_SCAN_THREADS) return;
    const int n = *count;
    for (int u = blockIdx.x; u < n; u += gridDim.x) {
        const int2 w = units[u];
        const PTK_GROUPBY_INST &I = insts[w.x];
        if (!gb_keys_ok(I.numKeys) || I.n < 0) continue;
        const PTK_GB_SCRATCH S = ptk_gb_scratch(I.scratch, I.n, I.numKeys);
        if (w.y < S.buckets) gb_bucketscan(S, w.y, red);
    }
}

// Units {inst, u} as in BatchGroupbyHistogram; unit 0 of an instance also writes its bucket and
// tile starts and tile table.  Each XCD takes a contiguous run of units (consecutive chunks of
// an instance): see GroupbyPartition.
extern "C" __global__ void __launch_bounds__(PTK_GB_PART_THREADS, 4)
BatchGroupbyPartition(const int2 *__restrict__ units, const int *__restrict__ count, const PTK_GROUPBY_INST *__restrict__ insts) {
    __shared__ GbPartLds L;
    if (blockDim.x != PTK_GB_PART_THREADS) return;
    const int n = *count, G = (int)gridDim.x < n ? (int)gridDim.x : n;
    if ((int)blockIdx.x >= G) return;
    for (int u = (int)gb_xcd_block(blockIdx.x, (unsigned)G); u < n; u += G) {
        const int2 w = units[u];
        const PTK_GROUPBY_INST &I = insts[w.x];
        if (!gb_keys_ok(I.numKeys) || I.n <= 0) continue;
        const PTK_GB_SCRATCH S = ptk_gb_scratch(I.scratch, I.n, I.numKeys);
        if (w.y == 0) gb_partition_starts(S, L);
        const long long per = I.chunkRows / PTK_GB_CHUNK, c0 = (long long)w.y * per;
        gb_partition_chunks(I.keys, I.vals, I.n, I.numKeys, c0, c0 + per < S.chunks ? c0 + per : S.chunks, 1, S, L);
    }
}

// Units {inst, tile} for every tile an instance may have (ceil(n / PTK_GB_TILE) + buckets); those
// past the instance's tiles do nothing.
extern "C" __global__ void __launch_bounds__(PTK_GB_COUNT_THREADS)
BatchGroupbyKeyCount(const int2 *__restrict__ units, const int *__restrict__ count, const PTK_GROUPBY_INST *__restrict__ insts) {
    __shared__ GbKeyCountLds L;
    if (blockDim.x != PTK_GB_COUNT_THREADS) return;
    const int n = *count;
    for (int u = blockIdx.x; u < n; u += gridDim.x) {
        const int2 w = units[u];
        const PTK_GROUPBY_INST &I = insts[w.x];
        if (!gb_keys_ok(I.numKeys) || I.n <= 0) continue;
        const PTK_GB_SCRATCH S = ptk_gb_scratch(I.scratch, I.n, I.numKeys);
        if ((unsigned)w.y < S.tstart[S.buckets]) gb_keycount_tile(S, (unsigned)w.y, L);
    }
}

// Units {inst, half bucket}.
extern "C" __global__ void __launch_bounds__(PTK_GB_KEYSCAN_THREADS)
BatchGroupbyKeyScan(const int2 *__restrict__ units, const int *__restrict__ count, const PTK_GROUPBY_INST *__restrict__ insts) {
    __shared__ GbKeyScanLds L;
    if (blockDim.x != PTK_GB_KEYSCAN_THREADS) return;
    const int n = *count;
    for (int u = blockIdx.x; u < n; u += gridDim.x) {
        const int2 w = units[u];
        const PTK_GROUPBY_INST &I = insts[w.x];
        if (!gb_keys_ok(I.numKeys) || I.n < 0) continue;
        const PTK_GB_SCRATCH S = ptk_gb_scratch(I.scratch, I.n, I.numKeys);
        if (w.y < 2 * S.buckets) gb_keyscan_half(S, I.numKeys, w.y / 2, w.y % 2, I.offsets, L);
    }
}

// Units {inst, tile} as in BatchGroupbyKeyCount, each XCD a contiguous run (see GroupbyReshuffle).
extern "C" __global__ void __launch_bounds__(PTK_GB_TILE_THREADS)
BatchGroupbyReshuffle(const int2 *__restrict__ units, const int *__restrict__ count, const PTK_GROUPBY_INST *__restrict__ insts) {
    __shared__ GbShuffleLds L;
    if (blockDim.x != PTK_GB_TILE_THREADS) return;
    const int n = *count, G = (int)gridDim.x < n ? (int)gridDim.x : n;
    if ((int)blockIdx.x >= G) return;
    for (int u = (int)gb_xcd_block(blockIdx.x, (unsigned)G); u < n; u += G) {
        const int2 w = units[u];
        const PTK_GROUPBY_INST &I = insts[w.x];
        if (!gb_keys_ok(I.numKeys) || I.n <= 0) continue;
        const PTK_GB_SCRATCH S = ptk_gb_scratch(I.scratch, I.n, I.numKeys);
        const unsigned t = (unsigned)w.y;
        if (t < S.tstart[S.buckets]) gb_reshuffle_tiles(S, I.numKeys, t, t + 1, 1, I.offsets, I.outKeys, I.outVals, I.n, L);
    }
}

// Units {inst, quarter bucket}.
extern "C" __global__ void __launch_bounds__(PTK_GB_TOTALS_THREADS)
BatchGroupbyTotals(const int2 *__restrict__ units, const int *__restrict__ count, const PTK_GROUPBY_INST *__restrict__ insts) {
    __shared__ GbTotalsLds L;
    if (blockDim.x != PTK_GB_TOTALS_THREADS) return;
    const int n = *count;
    for (int u = blockIdx.x; u < n; u += gridDim.x) {
        const int2 w = units[u];
        const PTK_GROUPBY_INST &I = insts[w.x];
        if (!gb_keys_ok(I.numKeys) || I.n < 0) continue;
        const PTK_GB_SCRATCH S = ptk_gb_scratch(I.scratch, I.n, I.numKeys);
        if (w.y < 4 * S.buckets) gb_totals_quarter(S, I.numKeys, w.y / 4, w.y % 4, I.table, I.groups, L);
    }
}

// Keys Zipf(s) by inverse CDF over nk = numKeys + 1 ranks, rank r -> key r*A mod nk (A coprime
// to nk), or uniform on [0, numKeys) when cdf is null; values are 32 random bits.  The same
// draws as the groupby graph application's host generator (streams 1 and 2 of the seed).
extern "C" __global__ void __launch_bounds__(256)
BatchInitGroupby(int *__restrict__ keys, int *__restrict__ vals, long long n, const double *__restrict__ cdf,
                 int numKeys, long long A, unsigned long long seed) {
    const long long nk = (long long)numKeys + 1;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const double u = ptk_rand_unit_f64(seed, 1, (uint64_t)i);
        long long k;
        if (cdf) {
            long long lo = 0, hi = nk;  // first rank whose cdf exceeds u (upper bound)
            while (lo < hi) {
                const long long mid = (lo + hi) >> 1;
                if (cdf[mid] <= u) lo = mid + 1;
                else hi = mid;
            }
            const long long r = lo >= nk ? nk - 1 : lo;
            k = (r * A) % nk;
        } else {
            k = (long long)(u * (double)numKeys);
        }
        keys[i] = (int)k;
        vals[i] = (int)(unsigned)ptk_rand_u64(seed, 2, (uint64_t)i);
    }
}

// ---- GMM --------------------------------------------------------------------------------------

extern "C" __global__ void __launch_bounds__(256)
BatchGmmAccumulate(const int2 *__restrict__ units, const int *__restrict__ count, const PTK_GMM_INST *__restrict__ insts) {
    __shared__ __align__(16) GmmShared sh;
    const int n = *count;
    for (int u = blockIdx.x; u < n; u += gridDim.x) {
        const int2 w = units[u];
        const PTK_GMM_INST &I = insts[w.x];
        gmm_accumulate_unit(I.X, I.d, I.model, I.partials, w.y, sh);
    }
}

// One work unit, called rather than inlined: inlined into the work-list loop, values the loop
// keeps live push the tile bodies (250 of 256 registers on their own) into spills inside the
// tile loop; called, the only scratch traffic is the callee-saved registers at entry and exit,
// once per unit.  The instance's dimensions and pointers are the block's: scalar registers.
__device__ __attribute__((noinline)) void gmm_mfma_unit_call(const PTK_GMM_INST *Ip, int b, GmmMfmaShared &sh) {
    const PTK_GMM_INST &I = *Ip;
    GMM_DIMS d;
    d.N = __builtin_amdgcn_readfirstlane(I.d.N);
    d.D = __builtin_amdgcn_readfirstlane(I.d.D);
    d.K = __builtin_amdgcn_readfirstlane(I.d.K);
    d.blocks = __builtin_amdgcn_readfirstlane(I.d.blocks);
    d.maxIter = __builtin_amdgcn_readfirstlane(I.d.maxIter);
    d.pad[0] = d.pad[1] = d.pad[2] = 0;
    // readfirstlane is 32-bit: a pointer goes through it in two halves
    auto uni = [](const void *p) {
        const unsigned long long v = (unsigned long long)p;
        const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
        const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
        return ((unsigned long long)hi << 32) | lo;
    };
    const float *X = (const float *)uni(I.X);
    const double *model = (const double *)uni(I.model);
    double *partials = (double *)uni(I.partials);
    gmm_mfma_unit(X, d, model, partials, __builtin_amdgcn_readfirstlane(b), sh);
}

// The same units with the products on the matrix cores (gmm_mfma.h), as the executor's plain
// path and the GMM graph launch GmmAccumulateMfma: two blocks per CU by their LDS.
extern "C" __global__ void __launch_bounds__(256, 2)
BatchGmmAccumulateMfma(const int2 *__restrict__ units, const int *__restrict__ count,
                       const PTK_GMM_INST *__restrict__ insts) {
    __shared__ __align__(16) GmmMfmaShared sh;
    const int n = *count;
    for (int u = blockIdx.x; u < n; u += gridDim.x) {
        const int2 w = units[u];
        const PTK_GMM_INST &I = insts[w.x];
        gmm_mfma_unit_call(&I, w.y, sh);
    }
}

extern "C" __global__ void __launch_bounds__(256)
BatchGmmFinalize(const int2 *__restrict__ units, const int *__restrict__ count, const PTK_GMM_INST *__restrict__ insts) {
    const int n = *count;
    for (int u = blockIdx.x; u < n; u += gridDim.x) {
        const int2 w = units[u];
        const PTK_GMM_INST &I = insts[w.x];
        gmm_finalize_block(I.model, I.partials, I.d, w.y);
    }
}

// Points of a seeded mixture: component c of point p uniform on [0, K) (stream 12); means
// U[-10,10] (stream 10) and standard deviations U[0.5,2] (stream 11) per (c, m); coordinates
// mean + sd * z, z by Box-Muller in binary32 from streams 13 and 14 (u1 in (0, 1]).  first[c] receives the smallest point
// index of component c (first must hold INT_MAX on entry).
extern "C" __global__ void __launch_bounds__(256)
BatchInitGmm(float *__restrict__ X, int N, int D, int K, unsigned long long seed, int *first) {
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < N; p += (long long)gridDim.x * blockDim.x) {
        int c = (int)(ptk_rand_unit_f64(seed, 12, (uint64_t)p) * K);
        if (c >= K) c = K - 1;
        atomicMin(&first[c], (int)p);
        for (int m = 0; m < D; ++m) {
            const uint64_t cm = (uint64_t)c * D + m, i = (uint64_t)p * D + m;
            const float mean = -10.0f + 20.0f * ptk_rand_unit_f32(seed, 10, cm);
            const float sd = 0.5f + 1.5f * ptk_rand_unit_f32(seed, 11, cm);
            const float u1 = 1.0f - ptk_rand_unit_f32(seed, 13, i), u2 = ptk_rand_unit_f32(seed, 14, i);
            X[i] = mean + sd * (sqrtf(-2.0f * __logf(u1)) * __cosf(6.2831853f * u2));
        }
    }
}

// The initial model of Gaumixmod's constructor (gaumixmod.h:14-27): frac 1/K, var 1e-10, means
// the first point of each generating component (point c*N/K if a component drew none), and the
// binary32 E-step table.  One block of 256.
extern "C" __global__ void __launch_bounds__(256)
BatchInitGmmModel(double *__restrict__ model, GMM_DIMS d, const float *__restrict__ X, const int *__restrict__ first) {
    const int D = d.D, K = d.K, Dp = ptk_gmm_dpad(&d), tid = threadIdx.x;
    const long long elems = ptk_gmm_model_elems(&d);
    for (long long i = tid; i < elems; i += blockDim.x) model[i] = 0.0;
    __syncthreads();
    if (tid == 0) {
        model[1] = d.N;
        model[2] = D;
        model[3] = K;
    }
    float *tab = (float *)(model + ptk_gmm_table(&d));
    for (int i = tid; i < K * D; i += blockDim.x) {
        const int c = i / D, m = i - c * D;
        const long long p = first[c] < d.N ? (long long)first[c] : (long long)c * d.N / K;
        const double mu = (double)X[p * D + m];
        model[ptk_gmm_mu(&d) + i] = mu;
        model[ptk_gmm_var(&d) + i] = 1.0e-10;
        tab[2 * c * Dp + m] = (float)mu;
        tab[2 * c * Dp + Dp + m] = (float)(1.0 / 1.0e-10);
    }
    for (int c = tid; c < K; c += blockDim.x) {
        model[ptk_gmm_frac(&d) + c] = 1.0 / K;
        tab[2 * K * Dp + c] = (float)(-0.5 * (D * log(1.0e-10)) + log(1.0 / K));
    }
}

// ---- checksums --------------------------------------------------------------------------------

__device__ inline void csum_add(unsigned long long *out, unsigned long long h) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) h += __shfl_xor(h, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, h);
}

// sum over i of mix(bits(a[i]), i): changes with any bit and any position of the array.
extern "C" __global__ void __launch_bounds__(256)
BatchChecksumWords(const unsigned *__restrict__ a, long long n, unsigned long long *out) {
    unsigned long long h = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        h += ptk_splitmix64(((unsigned long long)a[i] << 32) ^ (unsigned long long)i * 0x9E3779B97F4A7C15ull);
    csum_add(out, h);
}

// All checksums of a set of retiring instances in one launch: grid (stripes, jobs); blocks of
// job y stride over its elements.  Mode 2 (a groupby groups block) sums the header and the
// G = groups[0] triples it holds; the rest of the block is never written and is not part of the
// result.
extern "C" __global__ void __launch_bounds__(256)
BatchChecksumJobs(const PTK_CSUM_JOB *__restrict__ jobs) {
    const PTK_CSUM_JOB J = jobs[blockIdx.y];
    const long long n = J.mode == 2 ? 2 * (3 + 3 * ((const long long *)J.a)[0]) : J.n;
    const unsigned *a = (const unsigned *)J.a;
    const int *k = (const int *)J.a, *v = (const int *)J.b;
    unsigned long long h = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        if (J.mode == 1)
            h += ptk_splitmix64(((unsigned long long)(unsigned)k[i] << 32) | (unsigned)v[i]);
        else
            h += ptk_splitmix64(((unsigned long long)a[i] << 32) ^ (unsigned long long)i * 0x9E3779B97F4A7C15ull);
    }
    csum_add(J.out, h);
}

// The groups block of a groupby result: its header and the G = groups[0] triples it holds (the
// rest of the block is never written and is not part of the result).
extern "C" __global__ void __launch_bounds__(256)
BatchChecksumGroups(const long long *__restrict__ groups, unsigned long long *out) {
    const long long n = 2 * (3 + 3 * groups[0]);
    const unsigned *a = (const unsigned *)groups;
    unsigned long long h = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        h += ptk_splitmix64(((unsigned long long)a[i] << 32) ^ (unsigned long long)i * 0x9E3779B97F4A7C15ull);
    csum_add(out, h);
}

// sum over rows of mix(key, value): the multiset of (key, value) pairs, independent of order.
extern "C" __global__ void __launch_bounds__(256)
BatchChecksumPairs(const int *__restrict__ k, const int *__restrict__ v, long long n, unsigned long long *out) {
    unsigned long long h = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        h += ptk_splitmix64(((unsigned long long)(unsigned)k[i] << 32) | (unsigned)v[i]);
    csum_add(out, h);
}
