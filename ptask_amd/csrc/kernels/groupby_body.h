// groupby_body.h -- block-level bodies of the groupby task kernels, shared by the single-graph
// kernels (groupby.hip) and the batched multi-instance kernels (batch.hip).  Layouts of the
// table, groups and scratch arrays: include/ptk_groupby.h.
//
// No pass uses a global atomic on the data.  Per-key sums are exact int64 (wrapping) additions,
// whose result does not depend on their order, and every position is computed in row order, so
// every output -- including the order of rows inside a group, which is input order -- is
// bit-reproducible.
//
//   accumulate  a block aggregates one chunk of PTK_GB_CHUNK rows in an LDS hash table of
//               GB_SLOTS = 2 x PTK_GB_CHUNK slots (a chunk has at most PTK_GB_CHUNK distinct
//               keys, so the table never fills and every probe sequence ends), then writes the
//               chunk's distinct keys with their count and sum, grouped by key bucket.
//   merge       a block owns one (bucket, merge group) unit: PTK_GB_BUCKET keys over
//               PTK_GB_GROUP consecutive chunks.  It stages the chunks' entries of the bucket
//               in LDS, 64 chunks at a time, as a per-key mask of the chunks holding the key;
//               the thread of a key walks the staged chunks in chunk order, so the base it
//               writes into each entry (the key's rows in the group's earlier chunks) follows
//               row order.  The unit's per-key totals go to gcnt / gsum.
//   totals      a block per bucket turns the groups' per-key totals into per-group bases (in
//               group order) and writes the bucket's counts, sums and totals.
//   offsets     a block per bucket: the bucket totals of the lower buckets give its starting row
//               and group index; a block scan over its keys gives offsets and group records.
//   reshuffle   a block re-reads one chunk: each distinct key's first output row is offsets[key]
//               + its group base + its merge base; a row's rank among the chunk's rows of its key
//               is counted in row order (per wave with ballots, then wave bases); rows are staged
//               in LDS in output order and stored in that order.
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

#define GB_MERGE_BATCH 16  // chunks a merge unit stages at once: one bit of a 64-bit mask each
#define GB_MERGE_LOADS 4   // entries a thread loads before it stages them

struct GbMergeLds {
    unsigned tab[GB_MERGE_BATCH][PTK_GB_BUCKET];  // per staged chunk and key: index of the key's
                                                  // entry in the chunk's bucket segment << 16 | count
    unsigned long long mask[PTK_GB_BUCKET];       // per key: the staged chunks holding it
    unsigned long long sum[PTK_GB_BUCKET];
    unsigned segS[PTK_GB_GROUP];                  // per chunk of the group: its bucket segment
    unsigned ebase[GB_MERGE_BATCH];               // per staged chunk: index of its segment's first entry
    unsigned pre[PTK_GB_GROUP + 1];               // and the exclusive scan of the segment sizes
    unsigned long long red[16];
};

// Merge unit u = r * S.groups + g (bucket r, merge group g); blockDim.x == PTK_GB_BUCKET, thread k
// owns key r * PTK_GB_BUCKET + k.  Writes each entry's base inside the group and the group's
// per-key totals (gcnt, gsum).  The entries of a batch of chunks are loaded by all threads at
// once (flat over the batch's segments), so a unit waits for memory about once per batch; a
// thread then walks the staged chunks holding its key in chunk order.
__device__ inline void gb_merge_unit(const PTK_GB_SCRATCH &S, long long u, GbMergeLds &L) {
    const int NB = (int)S.buckets;
    const long long NG = S.groups;
    const int r = (int)(u / NG);
    const long long g = u % NG;
    const int k = threadIdx.x, lane = k & 63, w = k >> 6;
    const unsigned key0 = (unsigned)r * PTK_GB_BUCKET;
    const long long KP = (long long)NB * PTK_GB_BUCKET;
    const long long cg0 = g * PTK_GB_GROUP;
    const int nc = (int)(cg0 + PTK_GB_GROUP < S.chunks ? PTK_GB_GROUP : S.chunks - cg0);
    // the group's bucket segments (thread k: chunk cg0 + k) and their flat positions
    unsigned len = 0;
    if (k < nc) {
        const unsigned *bo = S.boff + (cg0 + k) * (NB + 1) + r;
        const unsigned s0 = bo[0] < PTK_GB_CHUNK ? bo[0] : PTK_GB_CHUNK;
        const unsigned s1 = bo[1] < PTK_GB_CHUNK ? bo[1] : PTK_GB_CHUNK;
        L.segS[k] = s0;
        len = s1 > s0 ? s1 - s0 : 0;
        len = len < PTK_GB_BUCKET ? len : PTK_GB_BUCKET;  // a segment holds distinct keys of one bucket
    }
    const unsigned long long incl = wave_incl(len);
    L.sum[k] = 0;
    L.mask[k] = 0;
    __syncthreads();  // the previous unit is done with red
    if (lane == 63) L.red[w] = incl;
    __syncthreads();
    unsigned long long pre = incl - len;
    for (int q = 0; q < w; ++q) pre += L.red[q];
    L.pre[k] = (unsigned)pre;
    if (k == PTK_GB_BUCKET - 1) L.pre[PTK_GB_BUCKET] = (unsigned)(pre + len);
    __syncthreads();
    unsigned run = 0;  // rows of the key in the chunks walked so far
    for (int j0 = 0; j0 < nc; j0 += GB_MERGE_BATCH) {
        const int nb = nc - j0 < GB_MERGE_BATCH ? nc - j0 : GB_MERGE_BATCH;
        const unsigned F0 = L.pre[j0], F1 = L.pre[j0 + nb];
        for (unsigned f0 = F0 + k; f0 < F1; f0 += GB_MERGE_LOADS * PTK_GB_BUCKET) {
            PTK_GB_ENTRY e[GB_MERGE_LOADS];
            int jj[GB_MERGE_LOADS];
            unsigned ii[GB_MERGE_LOADS];
#pragma unroll
            for (int q = 0; q < GB_MERGE_LOADS; ++q) {
                const unsigned f = f0 + q * PTK_GB_BUCKET;
                jj[q] = -1;
                if (f < F1) {
                    int lo = j0, hi = j0 + nb;  // the last chunk whose segment starts at or before f
                    while (hi - lo > 1) {
                        const int mid = (lo + hi) >> 1;
                        if (L.pre[mid] <= f) lo = mid;
                        else hi = mid;
                    }
                    jj[q] = lo;
                    ii[q] = f - L.pre[lo];
                    e[q] = S.entry[(cg0 + lo) * PTK_GB_CHUNK + L.segS[lo] + ii[q]];
                }
            }
#pragma unroll
            for (int q = 0; q < GB_MERGE_LOADS; ++q) {
                if (jj[q] < 0) continue;
                const unsigned kk = (unsigned)e[q].key - key0;
                if (kk >= PTK_GB_BUCKET) continue;  // only for an inconsistent scratch
                const int j = jj[q] - j0;
                atomicOr(&L.mask[kk], 1ull << j);
                L.tab[j][kk] = (ii[q] << 16) | (e[q].count & 0xFFFFu);
                atomicAdd(&L.sum[kk], (unsigned long long)e[q].sum);
            }
        }
        if (k < nb) L.ebase[k] = (unsigned)((cg0 + j0 + k) * PTK_GB_CHUNK + L.segS[j0 + k]);
        __syncthreads();
        unsigned long long m = L.mask[k];
        L.mask[k] = 0;
        while (m) {  // the key's chunks in chunk order, up to four per round (their reads overlap)
            int j[4];
            unsigned v[4], eb[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                j[q] = m ? __builtin_ctzll(m) : -1;
                m &= m - 1;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (j[q] >= 0) {
                    v[q] = L.tab[j[q]][k];
                    eb[q] = L.ebase[j[q]];
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (j[q] >= 0) {
                    S.entry[eb[q] + (v[q] >> 16)].sum = (long long)run;
                    run += v[q] & 0xFFFFu;
                }
            }
        }
        __syncthreads();
    }
    S.gsum[g * KP + key0 + k] = L.sum[k];
    S.gcnt[g * KP + key0 + k] = run;
}

// Bucket r (blockDim.x == PTK_GB_BUCKET, thread k owns key r * PTK_GB_BUCKET + k), after every
// merge unit: turns the groups' per-key totals into per-group bases in group order, writes the
// keys' counts and sums to the table, the bucket's rows and non-empty keys to S.btot and (bucket
// 0) the error count.
__device__ inline void gb_totals_bucket(const PTK_GB_SCRATCH &S, int numKeys, int r, long long *table,
                                        unsigned long long *red) {
    const int NB = (int)S.buckets, k = threadIdx.x;
    const long long NG = S.groups, KP = (long long)NB * PTK_GB_BUCKET;
    const long long key = (long long)r * PTK_GB_BUCKET + k;
    unsigned long long tot = 0, sum = 0;
    constexpr int kT = 8;  // groups whose totals are loaded together
    for (long long q0 = 0; q0 < NG; q0 += kT) {
        unsigned c[kT];
        unsigned long long v[kT];
#pragma unroll
        for (int q = 0; q < kT; ++q) {
            c[q] = q0 + q < NG ? S.gcnt[(q0 + q) * KP + key] : 0u;
            v[q] = q0 + q < NG ? S.gsum[(q0 + q) * KP + key] : 0ull;
        }
#pragma unroll
        for (int q = 0; q < kT; ++q) {
            if (q0 + q < NG) S.gcnt[(q0 + q) * KP + key] = (unsigned)tot;
            tot += c[q];
            sum += v[q];
        }
    }
    const bool in = key <= numKeys;
    if (in) {
        gb_count(table, numKeys)[key] = tot;
        gb_sum(table, numKeys)[key] = sum;
    }
    const unsigned long long rows = block_sum(in ? tot : 0, red);
    const unsigned long long nz = block_sum(in && tot ? 1 : 0, red);
    if (k == 0) {
        S.btot[2 * r] = (long long)rows;
        S.btot[2 * r + 1] = (long long)nz;
    }
    if (r == 0) {
        unsigned long long err = 0;
        for (long long c = k; c < S.chunks; c += blockDim.x) err += S.berr[c];
        err = block_sum(err, red);
        if (k == 0) *gb_errors(table, numKeys) = err;
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

#define GB_PLACE_WAVES (PTK_GB_PLACE_THREADS / 64)
#define GB_PLACE_ROWS (PTK_GB_CHUNK / PTK_GB_PLACE_THREADS)  // rows per thread
#define GB_ENTRY_BITS 11                                     // entry indices < PTK_GB_CHUNK

#define GB_KEY_MASK 0x7FFFFu  // keys < PTK_GB_MAX_KEYS = 2^19; a hash slot holds key | entry << 19

struct GbPlaceLds {
    union {
        unsigned slot[GB_SLOTS];  // hash of the chunk's distinct keys: key | entry index << 19
        struct {                  // once the rows are ranked: the rows in output order
            int row[PTK_GB_CHUNK], key[PTK_GB_CHUNK], val[PTK_GB_CHUNK];
        } stage;
    } u;
    int out0[PTK_GB_CHUNK];                             // per entry: output row of its first row
    unsigned short start[PTK_GB_CHUNK];                 // per entry: chunk-local position of its first row
    unsigned short wcnt[GB_PLACE_WAVES][PTK_GB_CHUNK];  // per wave and entry: rows, then the wave's base
    unsigned long long red[16];
};

// Chunk c: places its rows at offsets[key] + group base + merge base + (rank among the chunk's
// rows of the key, in row order), so a group holds its rows in input order.  Wave w counts the
// rows [w, w + 1) * PTK_GB_CHUNK / GB_PLACE_WAVES of the chunk in row order: a row's rank in its
// wave is the wave's running count of its entry plus the lower lanes of the same step holding the
// same entry (ballots over the entry index bits); the lower waves' counts are added afterwards.
// The chunk's rows are staged in LDS in output order and stored in that order, so a key's run
// of rows is written with consecutive stores.  blockDim.x == PTK_GB_PLACE_THREADS; begins and
// ends with a barrier.
__device__ inline void gb_place_chunk(const int *__restrict__ keys, const int *__restrict__ vals, long long n,
                                      int numKeys, long long c, const long long *__restrict__ offsets,
                                      const PTK_GB_SCRATCH &S, int *__restrict__ outKeys, int *__restrict__ outVals,
                                      GbPlaceLds &L) {
    const long long r0 = c * PTK_GB_CHUNK, r1 = r0 + PTK_GB_CHUNK < n ? r0 + PTK_GB_CHUNK : n;
    const int NB = (int)S.buckets;
    const unsigned *gbase = S.gcnt + (c / PTK_GB_GROUP) * ((long long)NB * PTK_GB_BUCKET);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (int s = tid; s < GB_SLOTS; s += PTK_GB_PLACE_THREADS) L.u.slot[s] = GB_EMPTY;
    for (int i = tid; i < GB_PLACE_WAVES * PTK_GB_CHUNK / 2; i += PTK_GB_PLACE_THREADS)
        ((unsigned *)&L.wcnt[0][0])[i] = 0;
    // the rows of this thread (row order: step s of wave w covers 64 consecutive rows)
    const long long wr0 = r0 + (long long)w * (PTK_GB_CHUNK / GB_PLACE_WAVES);
    int kr[GB_PLACE_ROWS], vr[GB_PLACE_ROWS];
#pragma unroll
    for (int s = 0; s < GB_PLACE_ROWS; ++s) {
        const long long row = wr0 + 64 * s + lane;
        kr[s] = row < r1 ? keys[row] : -1;
        vr[s] = row < r1 ? vals[row] : 0;
    }
    unsigned D = S.boff[c * (NB + 1) + NB];  // distinct keys of the chunk (<= PTK_GB_CHUNK)
    D = D < PTK_GB_CHUNK ? D : PTK_GB_CHUNK;
    // chunk-local start of each entry's rows: exclusive scan of the counts in entry order, each
    // thread taking a run of consecutive entries
    const unsigned per = (D + PTK_GB_PLACE_THREADS - 1) / PTK_GB_PLACE_THREADS, i0 = tid * per;
    const PTK_GB_ENTRY *ent = S.entry + c * PTK_GB_CHUNK;
    unsigned long long mine = 0;
    for (unsigned i = i0; i < i0 + per && i < D; ++i) mine += ent[i].count;
    const unsigned long long incl = wave_incl(mine);
    __syncthreads();  // the slots and counters are cleared; red is free
    if (lane == 63) L.red[w] = incl;
    __syncthreads();
    unsigned long long start = incl - mine;
    for (int q = 0; q < w; ++q) start += L.red[q];
    for (unsigned i = i0; i < i0 + per && i < D; ++i) {
        const int k = ent[i].key;
        if (k >= 0 && k <= numKeys) {
            const long long out0 = offsets[k] + (long long)gbase[k] + ent[i].sum;  // sum holds the merge base
            unsigned h = gb_hash((unsigned)k) & (GB_SLOTS - 1);
            for (int p = 0; p < GB_SLOTS; ++p) {  // distinct keys: the first empty slot is ours
                if (atomicCAS(&L.u.slot[h], GB_EMPTY, (unsigned)k | (i << 19)) == GB_EMPTY) break;
                h = (h + 1) & (GB_SLOTS - 1);
            }
            L.start[i] = (unsigned short)(start < PTK_GB_CHUNK ? start : PTK_GB_CHUNK);
            L.out0[i] = (int)out0;
        }
        start += ent[i].count;
    }
    __syncthreads();
    unsigned er[GB_PLACE_ROWS], rk[GB_PLACE_ROWS];
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
    for (int s = 0; s < GB_PLACE_ROWS; ++s) {
        const int k = kr[s];
        unsigned e = 0;
        bool use = false;
        if (k >= 0 && k <= numKeys) {
            unsigned h = gb_hash((unsigned)k) & (GB_SLOTS - 1);
            for (int p = 0; p < GB_SLOTS; ++p) {
                const unsigned cur = L.u.slot[h];
                if (cur == GB_EMPTY) break;  // only for an inconsistent scratch
                if ((cur & GB_KEY_MASK) == (unsigned)k) {
                    e = cur >> 19;
                    use = true;
                    break;
                }
                h = (h + 1) & (GB_SLOTS - 1);
            }
        }
        // lanes of this step holding the same entry
        unsigned long long peers = __ballot(use);
#pragma unroll
        for (int b = 0; b < GB_ENTRY_BITS; ++b) {
            const bool bit = (e >> b) & 1u;
            const unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const unsigned rank = (unsigned)__popcll(peers & below);
        unsigned old = 0;
        if (use) old = L.wcnt[w][e];
        if (use && rank == 0) L.wcnt[w][e] = (unsigned short)(old + (unsigned)__popcll(peers));
        er[s] = use ? e : 0xFFFFFFFFu;
        rk[s] = old + rank;
    }
    __syncthreads();  // the hash slots are dead from here on: their memory stages the rows
    for (unsigned e = tid; e < D; e += PTK_GB_PLACE_THREADS) {  // the lower waves' rows of each entry
        unsigned b = 0;
#pragma unroll
        for (int q = 0; q < GB_PLACE_WAVES; ++q) {
            const unsigned t = L.wcnt[q][e];
            L.wcnt[q][e] = (unsigned short)b;
            b += t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < GB_PLACE_ROWS; ++s) {
        if (er[s] == 0xFFFFFFFFu) continue;
        const unsigned e = er[s];
        const unsigned d = L.wcnt[w][e] + rk[s];  // the row's rank among the chunk's rows of its key
        const unsigned pos = (unsigned)L.start[e] + d;
        if (pos < PTK_GB_CHUNK) {  // always, for a consistent scratch
            L.u.stage.row[pos] = L.out0[e] + (int)d;
            L.u.stage.key[pos] = kr[s];
            L.u.stage.val[pos] = vr[s];
        }
    }
    __syncthreads();
    unsigned long long rows = 0;  // valid rows of the chunk: every local position below is filled
    for (int q = 0; q < GB_PLACE_WAVES; ++q) rows += L.red[q];
    rows = rows < PTK_GB_CHUNK ? rows : PTK_GB_CHUNK;
    for (unsigned p = tid; p < rows; p += PTK_GB_PLACE_THREADS) {
        const int row = L.u.stage.row[p];
        if (row >= 0 && row < n) {  // always, for a consistent scratch
            outKeys[row] = L.u.stage.key[p];
            outVals[row] = L.u.stage.val[p];
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
