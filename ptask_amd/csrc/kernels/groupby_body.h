// groupby_body.h -- block-level bodies of the groupby task kernels, shared by the single-graph
// kernels (groupby.hip) and the batched multi-instance kernels (batch.hip).
//
// table layout (int64): count[numKeys+1] | sum[numKeys+1] | cursor[numKeys+1] | errors[1].
// groups layout (int64): header {G, rows, errors} then G triples {key, count, sum} in
// ascending key order.  offsets[k] = rows with a smaller key (exclusive scan of count).
#pragma once

#include <hip/hip_runtime.h>

#define GB_LDS_SLOTS 4096   // per-block pre-aggregation table (power of two)
#define GB_PROBES 8         // linear probes before a row falls back to global atomics
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
__device__ inline unsigned long long *gb_cursor(long long *t, int numKeys) { return (unsigned long long *)t + 2ll * (numKeys + 1); }
__device__ inline unsigned long long *gb_errors(long long *t, int numKeys) { return (unsigned long long *)t + 3ll * (numKeys + 1); }

// Zero table elements i = start, start + stride, ...
__device__ inline void gb_clear(long long *table, int numKeys, long long start, long long stride) {
    const long long n = 3ll * (numKeys + 1) + 1;
    for (long long i = start; i < n; i += stride) table[i] = 0;
}

// Aggregates rows [r0, r1) into table.  Rows are read once (8 B each).  The block keeps a small
// open-addressed table in LDS so the frequent keys of a skewed (Zipf) stream are aggregated with
// LDS atomics; a key that finds no slot within GB_PROBES probes goes straight to the global
// table.  The block then flushes its LDS table with one global atomic per slot and field.
// Called by every thread of the block; begins and ends with a barrier-separated LDS phase.
__device__ inline void gb_accumulate_rows(const int *__restrict__ keys, const int *__restrict__ vals,
                                          long long r0, long long r1, int numKeys, long long *table) {
    __shared__ unsigned sKey[GB_LDS_SLOTS];
    __shared__ unsigned sCnt[GB_LDS_SLOTS];
    __shared__ unsigned long long sSum[GB_LDS_SLOTS];
    for (int s = threadIdx.x; s < GB_LDS_SLOTS; s += blockDim.x) {
        sKey[s] = GB_EMPTY;
        sCnt[s] = 0;
        sSum[s] = 0;
    }
    __syncthreads();
    unsigned long long *gCnt = gb_count(table, numKeys), *gSum = gb_sum(table, numKeys);
    unsigned long long localErr = 0;
    for (long long r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
        const int k = keys[r];
        const unsigned long long v = (unsigned long long)(long long)vals[r];
        if (k < 0 || k > numKeys) {
            ++localErr;
            continue;
        }
        const unsigned uk = (unsigned)k;
        unsigned h = gb_hash(uk) & (GB_LDS_SLOTS - 1);
        bool done = false;
        for (int p = 0; p < GB_PROBES; ++p) {
            unsigned cur = sKey[h];
            if (cur == GB_EMPTY) cur = atomicCAS(&sKey[h], GB_EMPTY, uk);
            if (cur == GB_EMPTY || cur == uk) {
                atomicAdd(&sCnt[h], 1u);
                atomicAdd(&sSum[h], v);
                done = true;
                break;
            }
            h = (h + 1) & (GB_LDS_SLOTS - 1);
        }
        if (!done) {
            atomicAdd(&gCnt[uk], 1ull);
            atomicAdd(&gSum[uk], v);
        }
    }
    __syncthreads();
    for (int s = threadIdx.x; s < GB_LDS_SLOTS; s += blockDim.x) {
        const unsigned k = sKey[s];
        if (k == GB_EMPTY) continue;
        atomicAdd(&gCnt[k], (unsigned long long)sCnt[s]);
        atomicAdd(&gSum[k], sSum[s]);
    }
    if (localErr) atomicAdd(gb_errors(table, numKeys), localErr);
    __syncthreads();
}

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

// One block (blockDim.x a multiple of 64, at most 1024): offsets = exclusive scan of counts;
// groups = the non-empty keys in ascending order with their count and sum, after the header
// {G, rows, errors}.  Wave w owns a contiguous run of keys (a multiple of 64).  Pass 1: each wave
// totals its run (rows and non-empty keys) with coalesced loads, eight 64-key steps in flight
// per batch; the wave totals are scanned in LDS.  Pass 2: each wave walks its run again, 64 keys
// per step: a wave-wide inclusive scan gives the offsets, a ballot gives each non-empty key its
// rank among the group records.
__device__ inline void gb_offsets_block(long long *table, int numKeys, long long *offsets, long long *groups) {
    __shared__ unsigned long long sRows[16], sNz[16], sRowBase[16], sNzBase[16];
    const unsigned long long *cnt = gb_count(table, numKeys), *sum = gb_sum(table, numKeys);
    const long long K = (long long)numKeys + 1;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const long long per = ((K + nw - 1) / nw + 63) / 64 * 64;
    const long long lo = min(K, per * w), hi = min(K, lo + per);
    unsigned long long rows = 0, nz = 0;
    for (long long base = lo; base < hi; base += 64 * 8) {
        unsigned long long c[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const long long k = base + i * 64 + lane;
            c[i] = k < hi ? cnt[k] : 0ull;
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            rows += c[i];
            nz += c[i] ? 1ull : 0ull;
        }
    }
    rows = wave_sum(rows);
    nz = wave_sum(nz);
    if (lane == 0) {
        sRows[w] = rows;
        sNz[w] = nz;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long r = 0, g = 0;
        for (int i = 0; i < nw; ++i) {
            sRowBase[i] = r;
            sNzBase[i] = g;
            r += sRows[i];
            g += sNz[i];
        }
        groups[0] = (long long)g;
        groups[1] = (long long)r;
        groups[2] = (long long)*gb_errors(table, numKeys);
    }
    __syncthreads();
    unsigned long long carry = sRowBase[w], gcarry = sNzBase[w];
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;  // lanes < this lane
    for (long long base = lo; base < hi; base += 64 * 8) {
        unsigned long long c[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const long long k = base + i * 64 + lane;
            c[i] = k < hi ? cnt[k] : 0ull;
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const long long k = base + i * 64 + lane;
            const bool in = k < hi;
            const unsigned long long incl = wave_incl(c[i]);
            if (in) offsets[k] = (long long)(carry + incl - c[i]);
            carry += __shfl(incl, 63, 64);
            const bool has = in && c[i] != 0;
            const unsigned long long m = __ballot(has);
            if (has) {
                long long *g = groups + 3 + 3 * (long long)(gcarry + __popcll(m & below));
                g[0] = k;
                g[1] = (long long)c[i];
                g[2] = (long long)sum[k];
            }
            gcarry += (unsigned long long)__popcll(m);
        }
    }
    __syncthreads();
}

// Scatters rows [r0, r1) of an n-row input to offsets[key] + (rows of that key placed before
// it), so equal keys are contiguous and the keys output is exact; the order of values within a
// group is not fixed.  Pass 1 counts the chunk's rows per key in an LDS table (the same
// insertion rule as gb_accumulate_rows: a key either gets a slot on its first row, and then
// every later row finds it, or never does).  Each slot then reserves its rows' range with ONE
// global atomic on the key's cursor; pass 2 re-reads the chunk (from cache) and places rows with
// an LDS cursor.  Keys without a slot fall back to one global atomic per row.
__device__ inline void gb_reshuffle_rows(const int *__restrict__ keys, const int *__restrict__ vals, long long n,
                                         long long r0, long long r1, int numKeys,
                                         const long long *__restrict__ offsets, long long *table,
                                         int *__restrict__ outKeys, int *__restrict__ outVals) {
    __shared__ unsigned sKey[GB_LDS_SLOTS];
    __shared__ unsigned sCnt[GB_LDS_SLOTS];
    __shared__ unsigned long long sBase[GB_LDS_SLOTS];
    unsigned long long *cursor = gb_cursor(table, numKeys);
    for (int s = threadIdx.x; s < GB_LDS_SLOTS; s += blockDim.x) {
        sKey[s] = GB_EMPTY;
        sCnt[s] = 0;
    }
    __syncthreads();
    for (long long r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
        const int k = keys[r];
        if (k < 0 || k > numKeys) continue;
        const unsigned uk = (unsigned)k;
        unsigned h = gb_hash(uk) & (GB_LDS_SLOTS - 1);
        for (int p = 0; p < GB_PROBES; ++p) {
            unsigned cur = sKey[h];
            if (cur == GB_EMPTY) cur = atomicCAS(&sKey[h], GB_EMPTY, uk);
            if (cur == GB_EMPTY || cur == uk) {
                atomicAdd(&sCnt[h], 1u);
                break;
            }
            h = (h + 1) & (GB_LDS_SLOTS - 1);
        }
    }
    __syncthreads();
    for (int s = threadIdx.x; s < GB_LDS_SLOTS; s += blockDim.x) {
        const unsigned k = sKey[s];
        if (k == GB_EMPTY) continue;
        const unsigned long long first = atomicAdd(&cursor[k], (unsigned long long)sCnt[s]);
        sBase[s] = (unsigned long long)offsets[k] + first;
        sCnt[s] = 0;
    }
    __syncthreads();
    for (long long r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
        const int k = keys[r];
        if (k < 0 || k > numKeys) continue;
        const unsigned uk = (unsigned)k;
        unsigned h = gb_hash(uk) & (GB_LDS_SLOTS - 1);
        long long pos = -1;
        for (int p = 0; p < GB_PROBES; ++p) {
            const unsigned cur = sKey[h];
            if (cur == uk) {
                pos = (long long)(sBase[h] + atomicAdd(&sCnt[h], 1u));
                break;
            }
            if (cur == GB_EMPTY) break;
            h = (h + 1) & (GB_LDS_SLOTS - 1);
        }
        if (pos < 0) pos = offsets[k] + (long long)atomicAdd(&cursor[k], 1ull);
        if (pos < 0 || pos >= n) continue;  // only possible if the table is inconsistent
        outKeys[pos] = k;
        outVals[pos] = vals[r];
    }
    __syncthreads();
}

}  // namespace ptk_groupby
