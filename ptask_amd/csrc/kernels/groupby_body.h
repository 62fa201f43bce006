// groupby_body.h -- block-level bodies of the groupby task kernels, shared by the single-graph
// kernels (groupby.hip) and the batched multi-instance kernels (batch.hip).  Layouts of the
// table, groups and scratch arrays: include/ptk_groupby.h.
//
// No pass uses a global atomic.  Per-key sums are exact int64 (wrapping) additions, whose result
// does not depend on their order, so every output except the order of values inside one group
// is bit-reproducible.
//
//   accumulate  a block aggregates one chunk of PTK_GB_CHUNK rows in an LDS hash table of
//               GB_SLOTS = 2 x PTK_GB_CHUNK slots (a chunk has at most PTK_GB_CHUNK distinct
//               keys, so the table never fills and every probe sequence ends), then writes the
//               chunk's distinct keys with their count and sum, grouped by key bucket.
//   merge       a block owns one bucket of PTK_GB_BUCKET keys: it adds every chunk's entries of
//               that bucket in LDS -- the value an entry's count is added to is the number of
//               rows of its key in the chunks merged before it, i.e. the start of the chunk's
//               range inside the key's group, which replaces the entry's sum in place -- and
//               writes the bucket's count and sum.
//   offsets     a block per bucket: the bucket totals of the lower buckets give its starting row
//               and group index; a block scan over its keys gives offsets and group records.
//   reshuffle   a block re-reads one chunk: each distinct key's first output row is
//               offsets[key] + its merge base; rows are staged in LDS in output order with LDS
//               cursors and stored in that order.
#pragma once

#include <hip/hip_runtime.h>

#include "ptk_groupby.h"

#define GB_SLOTS (2 * PTK_GB_CHUNK)  // LDS hash slots per chunk (power of two)
#define GB_EMPTY 0xFFFFFFFFu

namespace ptk_groupby {

__device__ inline unsigned gb_hash(unsigned k) {
    k ^= k >> 16;
    k *= 0x7feb352du;
    k ^= k >> 15;
    return k;
}

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

// Bijective renumbering of block b of n: blocks are dealt to the 8 XCDs round-robin, and this
// gives each XCD a contiguous run of the numbering.
__device__ inline unsigned gb_xcd_block(unsigned b, unsigned n) {
    if (n < 8) return b;
    const unsigned xcd = b % 8, local = b / 8, per = n / 8, rem = n % 8;
    return xcd * per + (xcd < rem ? xcd : rem) + local;
}

// Calls f(key, val) for the rows [r0, r1) that this thread owns: 4 consecutive rows per 16-byte
// load when the chunk starts on a multiple of 4.
template <class F>
__device__ inline void gb_for_rows(const int *__restrict__ keys, const int *__restrict__ vals, long long r0,
                                   long long r1, F f) {
    for (long long r = r0 + 4ll * threadIdx.x; r < r1; r += 4ll * blockDim.x) {
        if (r + 4 <= r1 && (r & 3) == 0) {
            const int4 k = *(const int4 *)&keys[r];
            const int4 v = *(const int4 *)&vals[r];
            f(k.x, v.x);
            f(k.y, v.y);
            f(k.z, v.z);
            f(k.w, v.w);
        } else {
            for (long long q = r; q < r1 && q < r + 4; ++q) f(keys[q], vals[q]);
        }
    }
}

struct GbChunkLds {
    unsigned key[GB_SLOTS];
    unsigned cnt[GB_SLOTS];
    unsigned long long sum[GB_SLOTS];
    unsigned hist[PTK_GB_MAX_BUCKETS + 1];  // entries per bucket, then running positions
    unsigned err;
};

// Chunk c of n rows: its distinct keys with count and sum, grouped by bucket, into S.entry, the
// per-bucket starts into S.boff[c], the invalid keys into S.berr[c].  Called by every thread of
// the block (blockDim.x a multiple of 64); begins and ends with a barrier-separated LDS phase.
__device__ inline void gb_accumulate_chunk(const int *__restrict__ keys, const int *__restrict__ vals, long long n,
                                           int numKeys, long long c, const PTK_GB_SCRATCH &S, GbChunkLds &L) {
    const long long r0 = c * PTK_GB_CHUNK, r1 = r0 + PTK_GB_CHUNK < n ? r0 + PTK_GB_CHUNK : n;
    const int NB = (int)S.buckets;
    for (int s = threadIdx.x; s < GB_SLOTS; s += blockDim.x) {
        L.key[s] = GB_EMPTY;
        L.cnt[s] = 0;
        L.sum[s] = 0;
    }
    for (int b = threadIdx.x; b <= NB; b += blockDim.x) L.hist[b] = 0;
    if (threadIdx.x == 0) L.err = 0;
    __syncthreads();
    gb_for_rows(keys, vals, r0, r1, [&](int k, int v) {
        if (k < 0 || k > numKeys) {
            atomicAdd(&L.err, 1u);
            return;
        }
        const unsigned uk = (unsigned)k;
        unsigned h = gb_hash(uk) & (GB_SLOTS - 1);
        for (int p = 0; p < GB_SLOTS; ++p) {  // at most PTK_GB_CHUNK slots are taken: ends early
            unsigned cur = L.key[h];
            if (cur == GB_EMPTY) cur = atomicCAS(&L.key[h], GB_EMPTY, uk);
            if (cur == GB_EMPTY || cur == uk) {
                atomicAdd(&L.cnt[h], 1u);
                atomicAdd(&L.sum[h], (unsigned long long)(long long)v);
                return;
            }
            h = (h + 1) & (GB_SLOTS - 1);
        }
    });
    __syncthreads();
    for (int s = threadIdx.x; s < GB_SLOTS; s += blockDim.x)
        if (L.key[s] != GB_EMPTY) atomicAdd(&L.hist[L.key[s] / PTK_GB_BUCKET], 1u);
    __syncthreads();
    unsigned *boff = S.boff + c * (NB + 1);
    if (threadIdx.x < 64) {  // one wave: exclusive scan of the bucket histogram
        const int lane = threadIdx.x;
        unsigned long long carry = 0;
        for (int b0 = 0; b0 < NB; b0 += 64) {
            const int b = b0 + lane;
            const unsigned long long v = b < NB ? L.hist[b] : 0;
            const unsigned long long incl = wave_incl(v);
            if (b < NB) {
                L.hist[b] = (unsigned)(carry + incl - v);
                boff[b] = (unsigned)(carry + incl - v);
            }
            carry += __shfl(incl, 63, 64);
        }
        if (lane == 0) {
            boff[NB] = (unsigned)carry;
            S.berr[c] = L.err;
        }
    }
    __syncthreads();
    PTK_GB_ENTRY *out = S.entry + c * PTK_GB_CHUNK;
    for (int s = threadIdx.x; s < GB_SLOTS; s += blockDim.x) {
        const unsigned k = L.key[s];
        if (k == GB_EMPTY) continue;
        const unsigned pos = atomicAdd(&L.hist[k / PTK_GB_BUCKET], 1u);
        PTK_GB_ENTRY e;
        e.key = (int)k;
        e.count = L.cnt[s];
        e.sum = (long long)L.sum[s];
        out[pos] = e;
    }
    __syncthreads();
}

struct GbMergeLds {
    unsigned long long cnt[PTK_GB_BUCKET];
    unsigned long long sum[PTK_GB_BUCKET];
    unsigned long long red[16];
};

// Bucket r: adds every chunk's entries of the bucket (the running count an entry is added to is
// its base), writes count and sum of the bucket's keys to the table and the bucket totals to
// S.btot; bucket 0 also writes the error count.  blockDim.x a multiple of 64, at most 1024.
__device__ inline void gb_merge_bucket(const PTK_GB_SCRATCH &S, int numKeys, int r, long long *table, GbMergeLds &L) {
    const int NB = (int)S.buckets;
    for (int k = threadIdx.x; k < PTK_GB_BUCKET; k += blockDim.x) {
        L.cnt[k] = 0;
        L.sum[k] = 0;
    }
    __syncthreads();
    const unsigned key0 = (unsigned)r * PTK_GB_BUCKET;
    constexpr int kAhead = 4;  // chunks whose segment bounds are loaded together
    for (long long c0 = threadIdx.x; c0 < S.chunks; c0 += (long long)kAhead * blockDim.x) {
        unsigned s[kAhead], e[kAhead];
#pragma unroll
        for (int j = 0; j < kAhead; ++j) {
            const long long c = c0 + (long long)j * blockDim.x;
            s[j] = e[j] = 0;
            if (c < S.chunks) {
                s[j] = S.boff[c * (NB + 1) + r];
                e[j] = S.boff[c * (NB + 1) + r + 1];
            }
        }
#pragma unroll
        for (int j = 0; j < kAhead; ++j) {
            const long long c = c0 + (long long)j * blockDim.x;
            for (unsigned i = s[j]; i < e[j]; ++i) {
                const long long at = c * PTK_GB_CHUNK + i;
                const PTK_GB_ENTRY E = S.entry[at];
                const unsigned k = (unsigned)E.key - key0;  // < PTK_GB_BUCKET by construction
                if (k >= PTK_GB_BUCKET) continue;
                atomicAdd(&L.sum[k], (unsigned long long)E.sum);
                S.entry[at].sum = (long long)atomicAdd(&L.cnt[k], (unsigned long long)E.count);  // the base
            }
        }
    }
    __syncthreads();
    unsigned long long rows = 0, nz = 0;
    for (int k = threadIdx.x; k < PTK_GB_BUCKET; k += blockDim.x) {
        const long long key = (long long)key0 + k;
        if (key > numKeys) continue;
        gb_count(table, numKeys)[key] = L.cnt[k];
        gb_sum(table, numKeys)[key] = L.sum[k];
        rows += L.cnt[k];
        nz += L.cnt[k] ? 1 : 0;
    }
    rows = block_sum(rows, L.red);
    nz = block_sum(nz, L.red);
    if (threadIdx.x == 0) {
        S.btot[2 * r] = (long long)rows;
        S.btot[2 * r + 1] = (long long)nz;
    }
    if (r == 0) {
        unsigned long long err = 0;
        for (long long c = threadIdx.x; c < S.chunks; c += blockDim.x) err += S.berr[c];
        err = block_sum(err, L.red);
        if (threadIdx.x == 0) *gb_errors(table, numKeys) = err;
    }
    __syncthreads();
}

struct GbOffsetsLds {
    unsigned long long red[16];
    unsigned long long wrows[16], wnz[16];
};

// Bucket r (blockDim.x == PTK_GB_BUCKET): offsets of its keys and their group records; bucket 0
// also writes the groups header {G, rows, errors}.
__device__ inline void gb_offsets_bucket(long long *table, int numKeys, const PTK_GB_SCRATCH &S, int r,
                                         long long *offsets, long long *groups, GbOffsetsLds &L) {
    const int NB = (int)S.buckets;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    unsigned long long rowsBefore = 0, nzBefore = 0, rowsAll = 0, nzAll = 0;
    for (int q = threadIdx.x; q < NB; q += blockDim.x) {
        const unsigned long long a = (unsigned long long)S.btot[2 * q], b = (unsigned long long)S.btot[2 * q + 1];
        if (q < r) {
            rowsBefore += a;
            nzBefore += b;
        }
        rowsAll += a;
        nzAll += b;
    }
    rowsBefore = block_sum(rowsBefore, L.red);
    nzBefore = block_sum(nzBefore, L.red);
    if (r == 0) {
        rowsAll = block_sum(rowsAll, L.red);
        nzAll = block_sum(nzAll, L.red);
        if (threadIdx.x == 0) {
            groups[0] = (long long)nzAll;
            groups[1] = (long long)rowsAll;
            groups[2] = (long long)*gb_errors(table, numKeys);
        }
    }
    const long long key = (long long)r * PTK_GB_BUCKET + threadIdx.x;
    const bool in = key <= numKeys;
    const unsigned long long c = in ? gb_count(table, numKeys)[key] : 0ull;
    const unsigned long long incl = wave_incl(c);
    const unsigned long long has = __ballot(c != 0);
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    if (lane == 63) {
        L.wrows[w] = incl;
        L.wnz[w] = (unsigned long long)__popcll(has);
    }
    __syncthreads();
    unsigned long long rowBase = rowsBefore, nzBase = nzBefore;
    for (int i = 0; i < w && i < nw; ++i) {
        rowBase += L.wrows[i];
        nzBase += L.wnz[i];
    }
    if (in) {
        offsets[key] = (long long)(rowBase + incl - c);
        if (c) {
            long long *g = groups + 3 + 3 * (long long)(nzBase + __popcll(has & below));
            g[0] = key;
            g[1] = (long long)c;
            g[2] = (long long)gb_sum(table, numKeys)[key];
        }
    }
    __syncthreads();
}

struct GbPlaceLds {
    unsigned key[GB_SLOTS];
    unsigned next[GB_SLOTS];   // next chunk-local position of the key's rows
    int delta[GB_SLOTS];       // output row - chunk-local position, for the key's rows
    int2 stage[PTK_GB_CHUNK];  // per chunk-local position: {output row, key}
    int sval[PTK_GB_CHUNK];    // and the value
    unsigned long long red[16];
};

// Chunk c: places its rows at offsets[key] + base(chunk, key) + (rank inside the chunk), so equal
// keys are contiguous and the keys output is exact; the order of values within a group is not
// fixed.  The chunk's rows are first staged in LDS in output order (each distinct key's rows
// contiguous), then stored in that order, so a key's run of rows is written with consecutive
// stores.  blockDim.x a multiple of 64; begins and ends with a barrier.
__device__ inline void gb_place_chunk(const int *__restrict__ keys, const int *__restrict__ vals, long long n,
                                      int numKeys, long long c, const long long *__restrict__ offsets,
                                      const PTK_GB_SCRATCH &S, int *__restrict__ outKeys, int *__restrict__ outVals,
                                      GbPlaceLds &L) {
    const long long r0 = c * PTK_GB_CHUNK, r1 = r0 + PTK_GB_CHUNK < n ? r0 + PTK_GB_CHUNK : n;
    const int NB = (int)S.buckets;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int s = threadIdx.x; s < GB_SLOTS; s += blockDim.x) L.key[s] = GB_EMPTY;
    unsigned D = S.boff[c * (NB + 1) + NB];  // distinct keys of the chunk (<= PTK_GB_CHUNK)
    D = D < PTK_GB_CHUNK ? D : PTK_GB_CHUNK;
    // chunk-local start of each entry's rows: exclusive scan of the counts in entry order, each
    // thread taking a run of consecutive entries
    const unsigned per = (D + blockDim.x - 1) / blockDim.x, i0 = threadIdx.x * per;
    unsigned long long mine = 0;
    for (unsigned i = i0; i < i0 + per && i < D; ++i) mine += S.entry[c * PTK_GB_CHUNK + i].count;
    const unsigned long long incl = wave_incl(mine);
    __syncthreads();  // the slots are cleared; red is free
    if (lane == 63) L.red[w] = incl;
    __syncthreads();
    unsigned long long start = incl - mine;
    for (int q = 0; q < w && q < nw; ++q) start += L.red[q];
    const PTK_GB_ENTRY *ent = S.entry + c * PTK_GB_CHUNK;
    for (unsigned i = i0; i < i0 + per && i < D; ++i) {
        const int k = ent[i].key;
        if (k >= 0 && k <= numKeys) {
            const long long out0 = offsets[k] + ent[i].sum;  // sum holds the merge base
            unsigned h = gb_hash((unsigned)k) & (GB_SLOTS - 1);
            for (int p = 0; p < GB_SLOTS; ++p) {  // distinct keys: the first empty slot is ours
                if (atomicCAS(&L.key[h], GB_EMPTY, (unsigned)k) == GB_EMPTY) {
                    L.next[h] = (unsigned)start;
                    L.delta[h] = (int)(out0 - (long long)start);
                    break;
                }
                h = (h + 1) & (GB_SLOTS - 1);
            }
        }
        start += ent[i].count;
    }
    __syncthreads();
    gb_for_rows(keys, vals, r0, r1, [&](int k, int v) {
        if (k < 0 || k > numKeys) return;
        unsigned h = gb_hash((unsigned)k) & (GB_SLOTS - 1);
        for (int p = 0; p < GB_SLOTS; ++p) {
            const unsigned cur = L.key[h];
            if (cur == (unsigned)k) {
                const unsigned lp = atomicAdd(&L.next[h], 1u);
                if (lp < PTK_GB_CHUNK) {  // always, for a consistent scratch
                    L.stage[lp] = make_int2(L.delta[h] + (int)lp, k);
                    L.sval[lp] = v;
                }
                return;
            }
            if (cur == GB_EMPTY) return;  // only for an inconsistent scratch
            h = (h + 1) & (GB_SLOTS - 1);
        }
    });
    __syncthreads();
    unsigned long long rows = 0;  // valid rows of the chunk: every local position below is filled
    for (int q = 0; q < nw; ++q) rows += L.red[q];
    rows = rows < PTK_GB_CHUNK ? rows : PTK_GB_CHUNK;
    for (unsigned p = threadIdx.x; p < rows; p += blockDim.x) {
        const int2 st = L.stage[p];
        if (st.x >= 0 && st.x < n) {  // always, for a consistent scratch
            outKeys[st.x] = st.y;
            outVals[st.x] = L.sval[p];
        }
    }
    __syncthreads();
}

struct GbCoalesceLds {
    long long g0;
    unsigned long long wtot[16];
};

// The coalesced layout of HostTasks/HostTasks.cpp:131-197 for rows [i0, i0 + PTK_GB_CHUNK) of a
// grouped output: group g (in key order) becomes the words key, then its rows' values, so row i
// of group g lands at word i + g + 1 and the key of g at start(g) + g; meta[g] = count of g.
// The block finds the group of its first row by binary search over the group starts
// (offsets[key of g], ascending in g) and the rest from key changes (a block scan of
// boundaries).  blockDim.x a multiple of 64, at most 1024.
__device__ inline void gb_coalesce_chunk(const int *__restrict__ vals, const int *__restrict__ keys,
                                         const long long *__restrict__ groups, const long long *__restrict__ offsets,
                                         long long c, int *__restrict__ out, int *__restrict__ meta, GbCoalesceLds &L) {
    const long long G = groups[0], rows = groups[1];
    const long long i0 = c * PTK_GB_CHUNK, i1 = i0 + PTK_GB_CHUNK < rows ? i0 + PTK_GB_CHUNK : rows;
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
    // each thread takes PTK_GB_CHUNK / blockDim.x consecutive rows
    const int per = (PTK_GB_CHUNK + blockDim.x - 1) / blockDim.x;
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
