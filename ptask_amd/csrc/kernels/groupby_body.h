// groupby_body.h -- block-level bodies of the groupby task kernels, shared by the single-graph
// kernels (groupby.hip) and the batched multi-instance kernels (batch.hip).  Layouts of the
// table, groups and scratch arrays: include/ptk_groupby.h.
//
// A stable two-level radix partition of the rows by key.  Level 1 splits the rows into buckets
// of PTK_GB_BUCKET consecutive keys, level 2 orders each bucket by key.  Every position is
// computed in row order (ballot ranks inside a wave, wave bases inside a block, chunk or tile
// prefixes across blocks), so a group holds its rows in input order and every output is
// bit-reproducible; per-key sums are exact int64 (wrapping) additions.  No pass uses a global
// atomic or waits on another block.
//
//   histogram   a block counts one chunk of PTK_GB_CHUNK rows per bucket (and its invalid keys)
//   bucketscan  a block per bucket: the exclusive prefix of its chunk counts over the chunks,
//               in place, and the bucket's rows
//   partition   a block re-reads one chunk, ranks its rows by bucket in row order, stages them
//               in LDS in that order and writes each bucket's run of rows behind the earlier
//               chunks' runs of the bucket: the key's low byte and the value
//   keycount    a block per tile of PTK_GB_TILE partitioned rows (tiles never straddle a bucket)
//               counts its rows per key
//   keyscan     a block per bucket: the exclusive prefix of its tiles' key counts over the
//               tiles, in place; the keys' counts and offsets
//   reshuffle   a block re-reads one tile, ranks its rows by key in row order, stages them in
//               LDS in that order, writes each key's run behind the earlier tiles' rows of the
//               key and adds up its values per key
//   totals      a block per bucket: the keys' sums over the tiles, the table, the group records
#pragma once

#include <hip/hip_runtime.h>

#include "ptk_groupby.h"

namespace ptk_groupby {

__device__ inline unsigned long long *gb_count(long long *t, int numKeys) { (void)numKeys; return (unsigned long long *)t; }
__device__ inline unsigned long long *gb_sum(long long *t, int numKeys) { return (unsigned long long *)t + (numKeys + 1); }
__device__ inline unsigned long long *gb_errors(long long *t, int numKeys) { return (unsigned long long *)t + 3ll * (numKeys + 1); }

// TRUE when the key range fits the bucket layout of the scratch arrays.
__device__ inline bool gb_keys_ok(int numKeys) { return numKeys >= 0 && (long long)numKeys + 1 <= PTK_GB_MAX_KEYS; }

__device__ inline unsigned long long wave_incl(unsigned long long v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        unsigned long long u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
    }
    return v;
}

__device__ inline unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Sum of v over the block (blockDim.x a multiple of 64, at most 1024); every thread gets it.
__device__ inline unsigned long long block_sum(unsigned long long v, unsigned long long *red) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    v = wave_sum(v);
    __syncthreads();  // red may still be read from a previous call
    if (lane == 0) red[w] = v;
    __syncthreads();
    unsigned long long t = 0;
    for (int i = 0; i < nw; ++i) t += red[i];
    return t;
}

__device__ inline unsigned wave_incl32(unsigned v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
    }
    return v;
}

// Exclusive scan of a[0, len) in LDS, in place (each thread a run of consecutive elements; the
// total is below 2^32); returns the total to every thread.  blockDim.x a multiple of 64, at most
// 1024; begins and ends with a barrier.
__device__ inline unsigned gb_lds_exscan(unsigned *a, int len, unsigned *red) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, nw = blockDim.x >> 6;
    const int per = (len + (int)blockDim.x - 1) / (int)blockDim.x, i0 = tid * per;
    __syncthreads();
    unsigned mine = 0;
    for (int i = i0; i < i0 + per && i < len; ++i) mine += a[i];
    const unsigned incl = wave_incl32(mine);
    if (lane == 63) red[w] = incl;
    __syncthreads();
    unsigned run = incl - mine, tot = 0;
    for (int q = 0; q < nw; ++q) {
        if (q < w) run += red[q];
        tot += red[q];
    }
    for (int i = i0; i < i0 + per && i < len; ++i) {
        const unsigned v = a[i];
        a[i] = run;
        run += v;
    }
    __syncthreads();
    return tot;
}

// Bijective renumbering of block b of n: blocks are dealt to the 8 XCDs round-robin, and this
// gives each XCD a contiguous run of the numbering.
__device__ inline unsigned gb_xcd_block(unsigned b, unsigned n) {
    if (n < 8) return b;
    const unsigned xcd = b % 8, local = b / 8, per = n / 8, rem = n % 8;
    return xcd * per + (xcd < rem ? xcd : rem) + local;
}

// Lanes of the wave holding the same digit d (of BITS bits) as this lane, among the lanes with
// use set.
template <int BITS>
__device__ inline unsigned long long gb_match(unsigned d, bool use) {
    const unsigned long long u = __ballot(use);
    unsigned lo = (unsigned)u, hi = (unsigned)(u >> 32);
#pragma unroll
    for (int b = 0; b < BITS; ++b) {
        // t: all ones where the bit is set; a lane keeps the lanes whose bit equals its own
        const unsigned t = (unsigned)__builtin_amdgcn_sbfe((int)d, b, 1);
        const unsigned long long m = __ballot(t != 0u);
        lo &= ~((unsigned)m ^ t);
        hi &= ~((unsigned)(m >> 32) ^ t);
    }
    return ((unsigned long long)hi << 32) | lo;
}

#define GB_HIST_WAVES (PTK_GB_HIST_THREADS / 64)
#define GB_HIST_ROWS (PTK_GB_CHUNK / PTK_GB_HIST_THREADS)  // rows per thread (16)

struct GbHistLds {
    unsigned cnt[PTK_GB_MAX_BUCKETS];  // per bucket: rows of the chunk
    unsigned long long red[16];
};

// Loads rows [wr0 + 64 s + lane] (s < R) of a (wave-uniform wr0): out-of-range rows read the
// last row and are flagged by row >= n.  base + wr0 is uniform, so the loads take a scalar base
// and a 32-bit lane offset.
template <int R, class T>
__device__ inline void gb_load_rows(const T *__restrict__ a, long long wr0, long long n, T (&v)[R]) {
    if (wr0 >= n) {
#pragma unroll
        for (int s = 0; s < R; ++s) v[s] = T(0);
        return;
    }
    const T *b = a + wr0;
    const unsigned lane = threadIdx.x & 63;
    if (wr0 + 64 * R <= n) {  // every row in range: one lane offset, immediate step offsets
#pragma unroll
        for (int s = 0; s < R; ++s) v[s] = b[64u * s + lane];
        return;
    }
    const long long lim = n - 1 - wr0;
    const unsigned last = lim < 64 * R - 1 ? (unsigned)lim : 64u * R - 1u;
#pragma unroll
    for (int s = 0; s < R; ++s) {
        const unsigned i = 64u * s + lane;
        v[s] = b[i < last ? i : last];
    }
}

// Chunk c: its rows per bucket into S.hist[c], its invalid keys into S.berr[c].  Each thread
// counts 4 consecutive rows per 16-byte load with LDS atomics that return nothing.
// blockDim.x == PTK_GB_HIST_THREADS; begins and ends with a barrier.
__device__ inline void gb_histogram_chunk(const int *__restrict__ keys, long long n, int numKeys, long long c,
                                          const PTK_GB_SCRATCH &S, GbHistLds &L) {
    const int NB = (int)S.buckets, tid = threadIdx.x;
    const long long r0 = c * PTK_GB_CHUNK, nr = n - r0 < PTK_GB_CHUNK ? n - r0 : PTK_GB_CHUNK;
    __syncthreads();
    for (int b = tid; b < NB; b += PTK_GB_HIST_THREADS) L.cnt[b] = 0;
    int4 kv[GB_HIST_ROWS / 4];
    const int4 *k4 = (const int4 *)(keys + r0);  // r0 is a multiple of 4: 16-byte aligned
    unsigned inRange = GB_HIST_ROWS;
    if (nr == PTK_GB_CHUNK) {
#pragma unroll
        for (int q = 0; q < GB_HIST_ROWS / 4; ++q) kv[q] = k4[q * PTK_GB_HIST_THREADS + tid];
    } else {  // the last chunk: rows past n read as -1 and are not counted
        inRange = 0;
        for (int q = 0; q < GB_HIST_ROWS / 4; ++q) {
            const long long i = 4ll * (q * PTK_GB_HIST_THREADS + tid);
            kv[q].x = i < nr ? keys[r0 + i] : -1;
            kv[q].y = i + 1 < nr ? keys[r0 + i + 1] : -1;
            kv[q].z = i + 2 < nr ? keys[r0 + i + 2] : -1;
            kv[q].w = i + 3 < nr ? keys[r0 + i + 3] : -1;
            const long long left = nr - i;
            inRange += left <= 0 ? 0u : (left >= 4 ? 4u : (unsigned)left);
        }
    }
    __syncthreads();
    unsigned good = 0;
    auto add = [&](int k) {
        if (k >= 0 && k <= numKeys) {
            atomicAdd(&L.cnt[(unsigned)k / PTK_GB_BUCKET], 1u);
            ++good;
        }
    };
#pragma unroll
    for (int q = 0; q < GB_HIST_ROWS / 4; ++q) {
        add(kv[q].x);
        add(kv[q].y);
        add(kv[q].z);
        add(kv[q].w);
    }
    const unsigned err = inRange - good;
    __syncthreads();
    unsigned *h = S.hist + c * NB;
    for (int b = tid; b < NB; b += PTK_GB_HIST_THREADS) h[b] = L.cnt[b];
    const unsigned long long e = block_sum(err, L.red);
    if (tid == 0) S.berr[c] = (unsigned)e;
    __syncthreads();
}

// Bucket b (any blockDim.x, a multiple of 64): S.hist[.][b] becomes the exclusive prefix over
// the chunks, S.btot[b] the bucket's rows; bucket 0 also adds up the invalid keys.
__device__ inline void gb_bucketscan(const PTK_GB_SCRATCH &S, int b, unsigned *red) {
    const int NB = (int)S.buckets, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const long long C = S.chunks, per = (C + blockDim.x - 1) / blockDim.x, c0 = (long long)tid * per;
    unsigned mine = 0;
    for (long long c = c0; c < c0 + per && c < C; ++c) mine += S.hist[c * NB + b];
    const unsigned incl = wave_incl32(mine);
    __syncthreads();
    if (lane == 63) red[w] = incl;
    __syncthreads();
    unsigned run = incl - mine, tot = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) {
        if (q < w) run += red[q];
        tot += red[q];
    }
    for (long long c = c0; c < c0 + per && c < C; ++c) {
        const unsigned v = S.hist[c * NB + b];
        S.hist[c * NB + b] = run;
        run += v;
    }
    if (tid == 0) S.btot[b] = tot;
    __syncthreads();
}

#define GB_PART_WAVES (PTK_GB_PART_THREADS / 64)
#define GB_PART_ROWS (PTK_GB_SUB / PTK_GB_PART_THREADS)  // rows per thread (8)
#define GB_PART_SUBS (PTK_GB_CHUNK / PTK_GB_SUB)         // quarter chunks per chunk

struct GbPartLds {
    union {
        unsigned short wcnt[GB_PART_WAVES][PTK_GB_MAX_BUCKETS];  // per wave and bucket: rows, then the wave's base
        int2 stage[PTK_GB_SUB];  // once the rows are ranked: {key, value} in partitioned order
    } u;
    unsigned dest[PTK_GB_MAX_BUCKETS];  // per bucket: the partitioned row of the chunk's next row of the bucket
    unsigned cnt[PTK_GB_MAX_BUCKETS];   // per bucket: the quarter's rows
    unsigned base[PTK_GB_MAX_BUCKETS];  // per bucket: local start, then the partitioned row of
                                        // local position 0 of the bucket
    unsigned bst[PTK_GB_MAX_BUCKETS];   // bucket starts (for the block's lifetime)
    unsigned red[16];
    unsigned long long red64[16];
};

// After the bucket scan, by one block of the partition launch: the bucket starts (S.bstart) and
// the first tile of each bucket (S.tstart), which the later passes read.  Any blockDim.x (a
// multiple of 64); begins and ends with a barrier.
__device__ inline void gb_partition_starts(const PTK_GB_SCRATCH &S, GbPartLds &L) {
    const int NB = (int)S.buckets, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    unsigned *ts = (unsigned *)&L.u.wcnt[0][0];  // tile starts (NB + 1 <= PTK_GB_MAX_BUCKETS + 1 entries)
    __syncthreads();
    for (int b = tid; b < NB; b += blockDim.x) L.base[b] = S.btot[b];
    const unsigned rows = gb_lds_exscan(L.base, NB, L.red);
    for (int b = tid; b < NB; b += blockDim.x) S.bstart[b] = L.base[b];
    if (tid == 0) S.bstart[NB] = rows;
    if (w == 0) {  // tiles of PTK_GB_TILE rows per bucket: exclusive scan of their counts
        unsigned carry = 0;
        for (int b0 = 0; b0 < NB; b0 += 64) {
            const int b = b0 + lane;
            const unsigned v = b < NB ? (S.btot[b] + PTK_GB_TILE - 1u) / PTK_GB_TILE : 0u;
            const unsigned incl = wave_incl32(v);
            if (b < NB) ts[b] = S.tstart[b] = carry + incl - v;
            carry += __shfl(incl, 63, 64);
        }
        if (lane == 0) ts[NB] = S.tstart[NB] = carry;
    }
    __syncthreads();
    unsigned long long err = 0;  // invalid keys
    for (long long c = tid; c < S.chunks; c += blockDim.x) err += S.berr[c];
    err = block_sum(err, L.red64);
    if (tid == 0) S.head[0] = err;
    for (int b = tid; b < NB; b += blockDim.x) {  // the tile table: bucket, first and end row
        const unsigned r0 = L.base[b], end = r0 + S.btot[b], t0 = ts[b], t1 = ts[b + 1];
        for (unsigned t = t0; t < t1; ++t) {
            const unsigned a = r0 + (t - t0) * PTK_GB_TILE;
            *(uint4 *)(S.tinfo + 4ull * t) = make_uint4((unsigned)b, a, a + PTK_GB_TILE < end ? a + PTK_GB_TILE : end, 0u);
        }
    }
    __syncthreads();
}

#define GB_PART_HX ((PTK_GB_MAX_BUCKETS + PTK_GB_PART_THREADS - 1) / PTK_GB_PART_THREADS)

// Chunks c, c + step, ... below cend, after the bucket scan, by one block, a quarter chunk at a
// time: writes each row's low key byte and value to S.tlow / S.tval at bstart[bucket] + the
// earlier chunks' rows of the bucket + the row's rank among the chunk's rows of the bucket (row
// order).  The next quarter's rows are loaded while the current quarter's are written.
// blockDim.x == PTK_GB_PART_THREADS; begins and ends with a barrier.
template <int BITS>
__device__ inline void gb_partition_chunks_t(const int *__restrict__ keys, const int *__restrict__ vals, long long n,
                                             int numKeys, long long c, long long cend, long long step,
                                             const PTK_GB_SCRATCH &S, GbPartLds &L) {
    const int NB = (int)S.buckets;
    const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    __syncthreads();
    for (int b = tid; b < NB; b += PTK_GB_PART_THREADS) L.bst[b] = S.btot[b];
    gb_lds_exscan(L.bst, NB, L.red);  // bucket starts
    int kr[GB_PART_ROWS], vr[GB_PART_ROWS];
    auto fetch = [&](long long row0) {
        const long long wr0 = row0 + (long long)w * (PTK_GB_SUB / GB_PART_WAVES);
        gb_load_rows<GB_PART_ROWS>(keys, wr0, n, kr);
        gb_load_rows<GB_PART_ROWS>(vals, wr0, n, vr);
    };
    if (c < cend) fetch(c * PTK_GB_CHUNK);
    // one quarter chunk per iteration: (c, q), q = 0..GB_PART_SUBS - 1, then c += step
    for (int q = 0; c < cend; q = q + 1 < GB_PART_SUBS ? q + 1 : 0, c += q == 0 ? step : 0) {
        if (q == 0)
            for (int b = tid; b < NB; b += PTK_GB_PART_THREADS) L.dest[b] = L.bst[b] + S.hist[c * NB + b];
        {
            const long long row0 = c * PTK_GB_CHUNK + (long long)q * PTK_GB_SUB;
            const long long wr0 = row0 + (long long)w * (PTK_GB_SUB / GB_PART_WAVES);
            const int nin = wr0 >= n ? 0 : (n - wr0 < 64 * GB_PART_ROWS ? (int)(n - wr0) : 64 * GB_PART_ROWS);
            for (int v = 0; v < GB_PART_WAVES; ++v)
                for (int b = tid; b < NB; b += PTK_GB_PART_THREADS) L.u.wcnt[v][b] = 0;
            __syncthreads();
            // ranks in row order: wave w, step s covers 64 consecutive rows; rk: the rank among the
            // wave's rows of the bucket, then the row's position in the quarter's partitioned order
            const unsigned long long below = (1ull << lane) - 1ull;
            unsigned rk[GB_PART_ROWS];
#pragma unroll
            for (int s = 0; s < GB_PART_ROWS; ++s) {
                const bool use = 64 * s + lane < nin && kr[s] >= 0 && kr[s] <= numKeys;
                const unsigned d = use ? (unsigned)kr[s] / PTK_GB_BUCKET : 0u;
                const unsigned long long peers = gb_match<BITS>(d, use);
                unsigned old = 0;
                if (use) old = L.u.wcnt[w][d];
                if (use && (peers & below) == 0) L.u.wcnt[w][d] = (unsigned short)(old + __popcll(peers));
                rk[s] = use ? old + (unsigned)__popcll(peers & below) : 0xFFFFFFFFu;
                __builtin_amdgcn_sched_barrier(0);  // one step at a time: keeps the ballots' masks few
            }
            __syncthreads();
            for (int b = tid; b < NB; b += PTK_GB_PART_THREADS) {  // wave bases and the quarter's rows per bucket
                unsigned t = 0;
#pragma unroll
                for (int v = 0; v < GB_PART_WAVES; ++v) {
                    const unsigned x = L.u.wcnt[v][b];
                    L.u.wcnt[v][b] = (unsigned short)t;
                    t += x;
                }
                L.base[b] = t;
                L.cnt[b] = t;
            }
            const unsigned valid = gb_lds_exscan(L.base, NB, L.red);  // local bucket starts
#pragma unroll
            for (int s = 0; s < GB_PART_ROWS; ++s) {
                if (rk[s] == 0xFFFFFFFFu) continue;
                const unsigned d = (unsigned)kr[s] / PTK_GB_BUCKET;
                rk[s] += L.base[d] + L.u.wcnt[w][d];
            }
            __syncthreads();
            // per bucket: the partitioned row of local position 0, and the next quarter's start
            for (int b = tid; b < NB; b += PTK_GB_PART_THREADS) {
                L.base[b] = L.dest[b] - L.base[b];
                L.dest[b] += L.cnt[b];
            }
#pragma unroll
            for (int s = 0; s < GB_PART_ROWS; ++s)
                if (rk[s] < PTK_GB_SUB) L.u.stage[rk[s]] = make_int2(kr[s], vr[s]);  // every valid row (rk < PTK_GB_SUB)
            __syncthreads();
            if (q + 1 < GB_PART_SUBS) fetch(row0 + PTK_GB_SUB);  // in flight while this quarter is written
            else if (c + step < cend) fetch((c + step) * PTK_GB_CHUNK);
            __builtin_amdgcn_sched_barrier(0);
            const unsigned nv = valid < PTK_GB_SUB ? valid : PTK_GB_SUB;
#pragma unroll
            for (int h = 0; h < GB_PART_ROWS; h += 4) {  // four staged rows per thread at a time
                int2 kv[4];
                unsigned g[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) kv[i] = L.u.stage[tid + (h + i) * PTK_GB_PART_THREADS];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const unsigned b = (unsigned)kv[i].x / PTK_GB_BUCKET;
                    g[i] = b < (unsigned)NB ? L.base[b] : 0u;
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const unsigned p = tid + (h + i) * PTK_GB_PART_THREADS, qq = g[i] + p;
                    if (p < nv && qq < (unsigned)n) {  // qq < n always, for a consistent scratch
                        S.tlow[qq] = (unsigned char)(kv[i].x & (PTK_GB_BUCKET - 1));
                        S.tval[qq] = kv[i].y;
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            __syncthreads();
        }
    }
}

// The two bucket-number widths: up to 256 buckets (65536 keys) and up to PTK_GB_MAX_BUCKETS.
__device__ inline void gb_partition_chunks(const int *__restrict__ keys, const int *__restrict__ vals, long long n,
                                           int numKeys, long long c, long long cend, long long step,
                                           const PTK_GB_SCRATCH &S, GbPartLds &L) {
    if (S.buckets <= 256) gb_partition_chunks_t<8>(keys, vals, n, numKeys, c, cend, step, S, L);
    else gb_partition_chunks_t<11>(keys, vals, n, numKeys, c, cend, step, S, L);
}

// Tile t of the partitioned rows (t < S.tstart[NB]): its bucket and row range.
struct GbTile {
    int bucket;
    unsigned r0, r1;
};

__device__ inline GbTile gb_tile(const PTK_GB_SCRATCH &S, unsigned t) {
    const uint4 v = *(const uint4 *)(S.tinfo + 4ull * t);
    GbTile T;
    T.bucket = (int)v.x;
    T.r0 = v.y;
    T.r1 = v.z;
    return T;
}

struct GbKeyCountLds {
    unsigned cnt[PTK_GB_BUCKET];
};

// Tile t: its rows per key into S.c2[t].  Each thread counts 16 consecutive rows (one 16-byte
// load of low key bytes) and adds runs of equal keys at once.  Any blockDim.x (a multiple of 64)
// of at least PTK_GB_BUCKET threads (PTK_GB_COUNT_THREADS in the kernels); begins and ends with
// a barrier.
__device__ inline void gb_keycount_tile(const PTK_GB_SCRATCH &S, unsigned t, GbKeyCountLds &L) {
    const GbTile T = gb_tile(S, t);
    const int tid = threadIdx.x;
    __syncthreads();
    if (tid < PTK_GB_BUCKET) L.cnt[tid] = 0;
    __syncthreads();
    const unsigned a0 = T.r0 & ~15u;
    for (unsigned q = a0 + 16u * tid; q < T.r1; q += 16u * blockDim.x) {
        const uint4 v = *(const uint4 *)(S.tlow + q);
        const unsigned wds[4] = {v.x, v.y, v.z, v.w};
        unsigned prev = 0xFFFFFFFFu, run = 0;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const unsigned r = q + i;
            if (r < T.r0 || r >= T.r1) continue;
            const unsigned k = (wds[i >> 2] >> (8 * (i & 3))) & 0xFFu;
            if (k != prev) {
                if (run) atomicAdd(&L.cnt[prev], run);
                prev = k;
                run = 0;
            }
            ++run;
        }
        if (run) atomicAdd(&L.cnt[prev], run);
    }
    __syncthreads();
    if (tid < PTK_GB_BUCKET) S.c2[(unsigned long long)t * PTK_GB_BUCKET + tid] = L.cnt[tid];
    __syncthreads();
}

#define GB_KS_KEYS (PTK_GB_BUCKET / 2)                    // keys per keyscan block (half a bucket)
#define GB_KS_PARTS (PTK_GB_KEYSCAN_THREADS / GB_KS_KEYS)    // threads per key in keyscan
#ifndef GB_KS_BATCH
#define GB_KS_BATCH 40  // tiles whose counts a keyscan thread loads at once
#endif
#define GB_TS_KEYS (PTK_GB_BUCKET / 4)                    // keys per totals block (a quarter bucket)
#define GB_TS_PARTS (PTK_GB_TOTALS_THREADS / GB_TS_KEYS)     // threads per key in totals
static_assert(GB_KS_KEYS % 64 == 0 && GB_TS_KEYS % 64 == 0, "a keyscan or totals part is whole waves");

struct GbKeyScanLds {
    unsigned part[GB_KS_PARTS][GB_KS_KEYS];
    unsigned wrows[16], wnz[16];
};

// Half h of bucket b (blockDim.x == PTK_GB_KEYSCAN_THREADS; thread q * GB_KS_KEYS + k takes part q of
// the bucket's tiles for key b * PTK_GB_BUCKET + h * GB_KS_KEYS + k): S.c2 of those keys over the
// bucket's tiles becomes the exclusive prefix over the tiles, in place; each key's count goes to
// S.kcnt, its first output row to offsets, and the half's non-empty keys to S.nz[2 b + h].  The
// upper half's first row is the bucket's first row plus its rows less the upper half's own.
__device__ inline void gb_keyscan_half(const PTK_GB_SCRATCH &S, int numKeys, int b, int h, long long *offsets,
                                       GbKeyScanLds &L) {
    const int tid = threadIdx.x, k = tid % GB_KS_KEYS, lane = tid & 63, w = tid >> 6;
    // the part is the same for a whole wave (GB_KS_KEYS is a multiple of 64): its tile range is
    // scalar, so each tile's counts load from a scalar row address plus the lane's key offset
    const int q = __builtin_amdgcn_readfirstlane(tid / GB_KS_KEYS);
    const unsigned t0 = S.tstart[b], t1 = S.tstart[b + 1], nt = t1 - t0;
    const unsigned per = (nt + GB_KS_PARTS - 1) / GB_KS_PARTS;
    const unsigned p0 = t0 + q * per, p1 = p0 + per < t1 ? p0 + per : t1;
    const unsigned kk = (unsigned)(h * GB_KS_KEYS + k);  // key within the bucket
    constexpr int kB = GB_KS_BATCH;  // tiles whose counts are loaded together
    unsigned v[kB];
    unsigned mine = 0;
    auto cnt = [&](unsigned t) -> unsigned & { return S.c2[(unsigned long long)t * PTK_GB_BUCKET + kk]; };
    for (unsigned t = p0; t < p1; t += kB) {
#pragma unroll
        for (int i = 0; i < kB; ++i) v[i] = t + i < p1 ? cnt(t + i) : 0u;
#pragma unroll
        for (int i = 0; i < kB; ++i) mine += v[i];
    }
    __syncthreads();
    L.part[q][k] = mine;
    __syncthreads();
    unsigned run = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < GB_KS_PARTS; ++i) {
        if (i < q) run += L.part[i][k];
        tot += L.part[i][k];
    }
    if (p1 - p0 <= (unsigned)kB && p1 > p0) {  // the counts are still in v
#pragma unroll
        for (int i = 0; i < kB; ++i) {
            if (p0 + i < p1) cnt(p0 + i) = run;
            run += v[i];
        }
    } else {
        for (unsigned t = p0; t < p1; t += kB) {
#pragma unroll
            for (int i = 0; i < kB; ++i) v[i] = t + i < p1 ? cnt(t + i) : 0u;
#pragma unroll
            for (int i = 0; i < kB; ++i) {
                if (t + i < p1) cnt(t + i) = run;
                run += v[i];
            }
        }
    }
    // part 0 of each key (the first GB_KS_KEYS / 64 waves): count, offsets, non-empty keys
    const long long key = (long long)b * PTK_GB_BUCKET + kk;
    const bool in = key <= numKeys;
    const unsigned c = (q == 0 && in) ? tot : 0u;
    const unsigned incl = wave_incl32(c);
    const unsigned long long has = __ballot(c != 0);
    if (lane == 63) {
        L.wrows[w] = incl;
        L.wnz[w] = (unsigned)__popcll(has);
    }
    __syncthreads();
    if (q == 0) {
        unsigned before = 0, nz = 0, half = 0;
        for (int i = 0; i < GB_KS_KEYS / 64; ++i) {
            if (i < w) before += L.wrows[i];
            nz += L.wnz[i];
            half += L.wrows[i];
        }
        const unsigned lower = h == 0 ? 0u : S.btot[b] - half;  // the rows of the lower half
        S.kcnt[key] = c;
        if (in) offsets[key] = (long long)(S.bstart[b] + lower + before + incl - c);
        if (tid == 0) S.nz[2 * b + h] = nz;
    }
    __syncthreads();
}

#define GB_TILE_WAVES (PTK_GB_TILE_THREADS / 64)
#define GB_TILE_ROWS (PTK_GB_TILE / PTK_GB_TILE_THREADS)  // rows per thread (8)

struct GbShuffleLds {
    unsigned short wcnt[GB_TILE_WAVES][PTK_GB_BUCKET];  // per wave and key: rows, then the wave's base
    unsigned start[PTK_GB_BUCKET];                      // per key: local start
    int base[PTK_GB_BUCKET];                            // per key: output row of local position 0
    unsigned long long sum[PTK_GB_BUCKET];
    alignas(16) int val[PTK_GB_TILE];                   // the tile's rows in key order
    alignas(16) unsigned char low[PTK_GB_TILE];
    unsigned red[16];
};

// Tiles t, t + step, ... below tend, after the key scan, by one block: places each tile's rows
// at offsets[key] + the earlier tiles' rows of the key + the row's rank among the tile's rows of
// the key (row order) and writes the tile's per-key sums to S.s2[t].  The next tile's rows are
// loaded while the current tile's are written.  blockDim.x == PTK_GB_TILE_THREADS; begins and
// ends with a barrier.
__device__ inline void gb_reshuffle_tiles(const PTK_GB_SCRATCH &S, int numKeys, unsigned t, unsigned tend,
                                          unsigned step, const long long *__restrict__ offsets,
                                          int *__restrict__ outKeys, int *__restrict__ outVals, long long n,
                                          GbShuffleLds &L) {
    const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    unsigned char kb[GB_TILE_ROWS];
    int vr[GB_TILE_ROWS];
    auto fetch = [&](const GbTile &T) {
        const long long wr0 = (long long)T.r0 + (long long)w * (PTK_GB_TILE / GB_TILE_WAVES);
        gb_load_rows<GB_TILE_ROWS>(S.tlow, wr0, (long long)T.r1, kb);
        gb_load_rows<GB_TILE_ROWS>(S.tval, wr0, (long long)T.r1, vr);
    };
    if (t >= tend) return;
    GbTile T = gb_tile(S, t);
    fetch(T);
    for (;;) {
        const unsigned rows = T.r1 - T.r0;
        const long long wr0 = (long long)T.r0 + (long long)w * (PTK_GB_TILE / GB_TILE_WAVES);
        const int nin = wr0 >= T.r1 ? 0 : (T.r1 - wr0 < 64 * GB_TILE_ROWS ? (int)(T.r1 - wr0) : 64 * GB_TILE_ROWS);
        __syncthreads();
        for (int q = 0; q < GB_TILE_WAVES; ++q)
            for (int k = tid; k < PTK_GB_BUCKET; k += PTK_GB_TILE_THREADS) L.wcnt[q][k] = 0;
        for (int k = tid; k < PTK_GB_BUCKET; k += PTK_GB_TILE_THREADS) L.sum[k] = 0;
        __syncthreads();
        const unsigned long long below = (1ull << lane) - 1ull;
        unsigned rk[GB_TILE_ROWS];
#pragma unroll
        for (int s = 0; s < GB_TILE_ROWS; ++s) {
            const bool use = 64 * s + lane < nin;
            const unsigned d = kb[s];
            const unsigned long long peers = gb_match<8>(d, use);
            unsigned old = 0;
            if (use) old = L.wcnt[w][d];
            if (use && (peers & below) == 0) L.wcnt[w][d] = (unsigned short)(old + __popcll(peers));
            rk[s] = use ? old + (unsigned)__popcll(peers & below) : 0xFFFFFFFFu;
            __builtin_amdgcn_sched_barrier(0);  // one step at a time: keeps the ballots' masks few
        }
        __syncthreads();
        for (int k = tid; k < PTK_GB_BUCKET; k += PTK_GB_TILE_THREADS) {
            unsigned tt = 0;
#pragma unroll
            for (int q = 0; q < GB_TILE_WAVES; ++q) {
                const unsigned v = L.wcnt[q][k];
                L.wcnt[q][k] = (unsigned short)tt;
                tt += v;
            }
            L.start[k] = tt;
        }
        gb_lds_exscan(L.start, PTK_GB_BUCKET, L.red);
        const long long key0 = (long long)T.bucket * PTK_GB_BUCKET;
        const unsigned *c2 = S.c2 + (unsigned long long)t * PTK_GB_BUCKET;
        for (int k = tid; k < PTK_GB_BUCKET; k += PTK_GB_TILE_THREADS) {
            const long long o = key0 + k <= numKeys ? offsets[key0 + k] : 0;  // keys past numKeys have no rows
            L.base[k] = (int)(o + (long long)c2[k] - (long long)L.start[k]);
        }
#pragma unroll
        for (int s = 0; s < GB_TILE_ROWS; ++s) {
            if (rk[s] == 0xFFFFFFFFu) continue;
            const unsigned p = L.start[kb[s]] + L.wcnt[w][kb[s]] + rk[s];
            if (p < PTK_GB_TILE) {  // always
                L.val[p] = vr[s];
                L.low[p] = kb[s];
            }
        }
        __syncthreads();
        const unsigned tn = t + step;
        if (tn < tend) {  // in flight while this tile is written
            T = gb_tile(S, tn);
            fetch(T);
        }
        // per-key sums: each thread adds up the runs of its GB_TILE_ROWS consecutive staged rows
        if ((unsigned)tid * GB_TILE_ROWS < rows) {
            const unsigned p0 = (unsigned)tid * GB_TILE_ROWS;
            unsigned lw[GB_TILE_ROWS / 4];
            int v[GB_TILE_ROWS];
#pragma unroll
            for (int q = 0; q < GB_TILE_ROWS / 4; ++q) {
                lw[q] = *(const unsigned *)&L.low[p0 + 4 * q];
                const int4 x = *(const int4 *)&L.val[p0 + 4 * q];
                v[4 * q] = x.x;
                v[4 * q + 1] = x.y;
                v[4 * q + 2] = x.z;
                v[4 * q + 3] = x.w;
            }
            const unsigned m = rows - p0 < GB_TILE_ROWS ? rows - p0 : GB_TILE_ROWS;
            unsigned prev = lw[0] & 0xFFu;
            unsigned long long acc = 0;
#pragma unroll
            for (int i = 0; i < GB_TILE_ROWS; ++i) {
                if ((unsigned)i >= m) break;
                const unsigned k = (lw[i >> 2] >> (8 * (i & 3))) & 0xFFu;
                if (k != prev) {
                    atomicAdd(&L.sum[prev], acc);
                    prev = k;
                    acc = 0;
                }
                acc += (unsigned long long)(long long)v[i];
            }
            atomicAdd(&L.sum[prev], acc);
        }
        unsigned kk[GB_TILE_ROWS];
        int vv[GB_TILE_ROWS];
#pragma unroll
        for (int i = 0; i < GB_TILE_ROWS; ++i) {
            kk[i] = L.low[tid + i * PTK_GB_TILE_THREADS];
            vv[i] = L.val[tid + i * PTK_GB_TILE_THREADS];
        }
        int gb[GB_TILE_ROWS];
#pragma unroll
        for (int i = 0; i < GB_TILE_ROWS; ++i) gb[i] = L.base[kk[i]];
#pragma unroll
        for (int i = 0; i < GB_TILE_ROWS; ++i) {
            const unsigned p = tid + i * PTK_GB_TILE_THREADS, q = (unsigned)(gb[i] + (int)p);
            if (p < rows && q < (unsigned)n) {  // q < n always, for a consistent scratch
                outKeys[q] = (int)key0 + (int)kk[i];
                outVals[q] = vv[i];
            }
        }
        __syncthreads();
        unsigned long long *s2 = S.s2 + (unsigned long long)t * PTK_GB_BUCKET;
        for (int k = tid; k < PTK_GB_BUCKET; k += PTK_GB_TILE_THREADS) s2[k] = L.sum[k];
        if (tn >= tend) break;
        t = tn;
    }
    __syncthreads();
}

struct GbTotalsLds {
    unsigned long long part[GB_TS_PARTS][GB_TS_KEYS];
    unsigned long long red[16];
};

// Quarter r of bucket b (blockDim.x == PTK_GB_TOTALS_THREADS; thread q * GB_TS_KEYS + k takes part
// q of the bucket's tiles for key b * PTK_GB_BUCKET + r * GB_TS_KEYS + k), after the reshuffle:
// the keys' sums over the bucket's tiles (tile order), the table's counts and sums and the keys'
// group records; the last quarter of the last bucket also writes the error count and the groups
// header {G, rows, errors}.
__device__ inline void gb_totals_quarter(const PTK_GB_SCRATCH &S, int numKeys, int b, int r, long long *table,
                                         long long *groups, GbTotalsLds &L) {
    const int NB = (int)S.buckets;
    const int tid = threadIdx.x, k = tid % GB_TS_KEYS;
    const int q = __builtin_amdgcn_readfirstlane(tid / GB_TS_KEYS);  // one part per wave
    const unsigned t0 = S.tstart[b], t1 = S.tstart[b + 1], nt = t1 - t0;
    const unsigned per = (nt + GB_TS_PARTS - 1) / GB_TS_PARTS;
    const unsigned p0 = t0 + q * per, p1 = p0 + per < t1 ? p0 + per : t1;
    const unsigned kk = (unsigned)(r * GB_TS_KEYS + k);  // key within the bucket
    constexpr int kB = 24;
    unsigned long long mine = 0;
    for (unsigned t = p0; t < p1; t += kB) {
        unsigned long long v[kB];
#pragma unroll
        for (int i = 0; i < kB; ++i) v[i] = t + i < p1 ? S.s2[(unsigned long long)(t + i) * PTK_GB_BUCKET + kk] : 0ull;
#pragma unroll
        for (int i = 0; i < kB; ++i) mine += v[i];
    }
    // group records before this quarter's keys: the non-empty keys of the lower buckets, then of
    // this bucket's lower keys
    unsigned long long nzb = 0;
    for (int i = tid; i < 2 * b; i += blockDim.x) nzb += S.nz[i];
    for (int i = tid; i < r * GB_TS_KEYS; i += blockDim.x) nzb += S.kcnt[(long long)b * PTK_GB_BUCKET + i] ? 1u : 0u;
    nzb = block_sum(nzb, L.red);
    L.part[q][k] = mine;
    const long long key = (long long)b * PTK_GB_BUCKET + kk;
    const bool in = key <= numKeys;
    const unsigned c = (q == 0 && in) ? S.kcnt[key] : 0u;
    const unsigned long long has = __ballot(c != 0);  // wave 0 holds the quarter's GB_TS_KEYS keys
    __syncthreads();
    if (q == 0) {
        unsigned long long sum = 0;
#pragma unroll
        for (int i = 0; i < GB_TS_PARTS; ++i) sum += L.part[i][k];
        const unsigned long long g = nzb + (unsigned long long)__popcll(has & ((1ull << k) - 1ull));
        if (in) {
            gb_count(table, numKeys)[key] = c;
            gb_sum(table, numKeys)[key] = sum;
            if (c) {
                long long *e = groups + 3 + 3 * (long long)g;
                e[0] = key;
                e[1] = (long long)c;
                e[2] = (long long)sum;
            }
        }
        if (b == NB - 1 && r == 3 && tid == 0) {  // the groups header and the error count
            const unsigned long long errors = S.head[0];
            *gb_errors(table, numKeys) = errors;
            groups[0] = (long long)(nzb + (unsigned long long)__popcll(has));
            groups[1] = (long long)S.bstart[NB];
            groups[2] = (long long)errors;
        }
    }
    __syncthreads();
}

#define GB_COALESCE_ROWS 2048  // grouped rows per GroupbyCoalesce unit

struct GbCoalesceLds {
    long long g0;
    unsigned long long wtot[16];
};

// The coalesced layout of HostTasks/HostTasks.cpp:131-197 for rows [i0, i0 + GB_COALESCE_ROWS) of a
// grouped output: group g (in key order) becomes the words key, then its rows' values, so row i
// of group g lands at word i + g + 1 and the key of g at start(g) + g; meta[g] = count of g.
// The block finds the group of its first row by binary search over the group starts
// (offsets[key of g], ascending in g) and the rest from key changes (a block scan of
// boundaries).  blockDim.x a multiple of 64, at most 1024.
__device__ inline void gb_coalesce_chunk(const int *__restrict__ vals, const int *__restrict__ keys,
                                         const long long *__restrict__ groups, const long long *__restrict__ offsets,
                                         long long c, int *__restrict__ out, int *__restrict__ meta, GbCoalesceLds &L) {
    const long long G = groups[0], rows = groups[1];
    const long long i0 = c * GB_COALESCE_ROWS, i1 = i0 + GB_COALESCE_ROWS < rows ? i0 + GB_COALESCE_ROWS : rows;
    if (i0 >= i1 || G <= 0) return;  // uniform over the block
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (threadIdx.x == 0) {  // the last group whose start is <= i0
        long long lo = 0, hi = G;
        while (hi - lo > 1) {
            const long long mid = (lo + hi) / 2;
            if (offsets[groups[3 + 3 * mid]] <= i0) lo = mid;
            else hi = mid;
        }
        L.g0 = lo;
    }
    // each thread takes GB_COALESCE_ROWS / blockDim.x consecutive rows
    const int per = (GB_COALESCE_ROWS + blockDim.x - 1) / blockDim.x;
    const long long a = i0 + (long long)threadIdx.x * per, b = a + per < i1 ? a + per : i1;
    unsigned long long mine = 0;
    for (long long i = a; i < b; ++i) mine += (i > i0 && keys[i] != keys[i - 1]) ? 1 : 0;
    const unsigned long long incl = wave_incl(mine);
    if (lane == 63) L.wtot[w] = incl;
    __syncthreads();
    long long g = L.g0 + (long long)(incl - mine);
    for (int q = 0; q < w && q < nw; ++q) g += (long long)L.wtot[q];
    for (long long i = a; i < b; ++i) {
        const bool starts = i > i0 ? keys[i] != keys[i - 1] : offsets[groups[3 + 3 * g]] == i0;
        if (i > i0 && starts) ++g;
        if (g >= G) break;  // only for an inconsistent input
        if (starts) {
            out[i + g] = keys[i];
            meta[g] = (int)groups[3 + 3 * g + 1];
        }
        out[i + g + 1] = vals[i];
    }
    __syncthreads();
}

}  // namespace ptk_groupby
