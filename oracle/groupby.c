/*
 * oracle/groupby.c -- CPU restatements of the prefix sum and the groupby graph.  TEST
 * INFRASTRUCTURE ONLY (the checker; no product code links it).
 *
 * Prefix sum: host_pfxsum (ptaskUnitTest/pfxsum.cpp:56-69): out[0] = 0, out[i] = out[i-1] +
 * in[i-1], int32 with two's-complement wrap (computed in unsigned arithmetic).
 *
 * Groupby: the reference defines its output by an invariant (graphgroupby.cpp:394-419): the
 * output is the input's keys reordered so equal keys are contiguous.  This restatement fixes one
 * such order -- groups by ascending key, rows of a group in input order (a stable counting
 * sort) -- and the per-key int64 count and sum of an int32 value column.  Keys outside
 * [0, numKeys] are counted as errors and dropped.
 *
 * Host-task forms with the HIP kernels' ABI (ptask_amd/csrc/kernels/groupby.hip) let the same
 * graph run on the HostAccelerator.
 *
 * Pinned by tests/test_oracle_groupby.py: the g_data fixture of pfxsum.cpp:22-40 and the
 * cpfxsum.cpp fixture prefix, the closed form sum_{j<i}, wrap-around at INT_MAX; groupby against
 * a sort + dictionary formulation in numpy and the reference's contiguity/set invariant.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "ptk_groupby.h"
#include "ptk_oracle.h"

void ptk_oracle_pfxsum_i32(const int32_t *in, int32_t *out, long long n) {
    uint32_t acc = 0;
    for (long long i = 0; i < n; ++i) {
        out[i] = (int32_t)acc;
        acc += (uint32_t)in[i];
    }
}

void ptk_oracle_pfxsum_i64(const int64_t *in, int64_t *out, long long n) {
    uint64_t acc = 0;
    for (long long i = 0; i < n; ++i) {
        out[i] = (int64_t)acc;
        acc += (uint64_t)in[i];
    }
}

/* table layout as the kernels use it: count | sum | cursor (numKeys+1 each) | errors */
static void gb_clear(int64_t *table, int numKeys) {
    memset(table, 0, sizeof(int64_t) * (3 * ((size_t)numKeys + 1) + 1));
}

static void gb_accumulate(const int32_t *keys, const int32_t *vals, long long n, int numKeys, int64_t *table) {
    uint64_t *cnt = (uint64_t *)table, *sum = cnt + (numKeys + 1), *err = cnt + 3 * ((size_t)numKeys + 1);
    for (long long r = 0; r < n; ++r) {
        int32_t k = keys[r];
        if (k < 0 || k > numKeys) {
            ++*err;
            continue;
        }
        cnt[k] += 1;
        sum[k] += (uint64_t)(int64_t)vals[r];
    }
}

static void gb_offsets(const int64_t *table, int numKeys, int64_t *offsets, int64_t *groups) {
    const uint64_t *cnt = (const uint64_t *)table, *sum = cnt + (numKeys + 1);
    uint64_t rows = 0, g = 0;
    for (int k = 0; k <= numKeys; ++k) {
        offsets[k] = (int64_t)rows;
        if (cnt[k]) {
            int64_t *e = groups + 3 + 3 * g;
            e[0] = k;
            e[1] = (int64_t)cnt[k];
            e[2] = (int64_t)sum[k];
            ++g;
        }
        rows += cnt[k];
    }
    groups[0] = (int64_t)g;
    groups[1] = (int64_t)rows;
    groups[2] = table[3 * ((size_t)numKeys + 1)];
}

static void gb_reshuffle(const int32_t *keys, const int32_t *vals, long long n, int numKeys,
                         const int64_t *offsets, int64_t *table, int32_t *outKeys, int32_t *outVals) {
    uint64_t *cursor = (uint64_t *)table + 2 * ((size_t)numKeys + 1);
    for (long long r = 0; r < n; ++r) {
        int32_t k = keys[r];
        if (k < 0 || k > numKeys) continue;
        int64_t pos = offsets[k] + (int64_t)cursor[k]++;
        outKeys[pos] = k;
        outVals[pos] = vals[r];
    }
}

/* The whole pipeline; groups needs 3 + 3*(numKeys+1) int64, offsets numKeys+1. */
void ptk_oracle_groupby(const int32_t *keys, const int32_t *vals, long long n, int numKeys,
                        int64_t *groups, int64_t *offsets, int32_t *outKeys, int32_t *outVals) {
    int64_t *table = (int64_t *)malloc(sizeof(int64_t) * (3 * ((size_t)numKeys + 1) + 1));
    gb_clear(table, numKeys);
    gb_accumulate(keys, vals, n, numKeys, table);
    gb_offsets(table, numKeys, offsets, groups);
    gb_reshuffle(keys, vals, n, numKeys, offsets, table, outKeys, outVals);
    free(table);
}

/* ---- host-task ABI forms (void f(UINT, void**)); INT arguments arrive by value ------------- */
#define IARG(a, i) ((int)(intptr_t)(a)[i])

void pfxsum(unsigned int n, void **a) {
    (void)n;
    ptk_oracle_pfxsum_i32((const int32_t *)a[0], (int32_t *)a[1], IARG(a, 2));
}

/* The two-level radix partition of the HIP kernels (include/ptk_groupby.h) restated on the host
 * with the same scratch layout: each pass fills the arrays the device pass fills, with the same
 * values (counts, prefixes and positions are functions of the input in row order, so a
 * sequential walk over the rows gives what the device's ballot ranks give). */
#define GB_SCRATCH_OR_RETURN(S, sp, n, numKeys)                                                \
    if ((n) < 0 || (numKeys) < 0 || (long long)(numKeys) + 1 > PTK_GB_MAX_KEYS) return;           \
    PTK_GB_SCRATCH S = ptk_gb_scratch((sp), (n), (numKeys))

/* GroupbyHistogram(keys, n, numKeys, scratch): rows per bucket per chunk, invalid keys per chunk. */
void GroupbyHistogram(unsigned int n_, void **a) {
    (void)n_;
    const int32_t *keys = (const int32_t *)a[0];
    const long long n = IARG(a, 1);
    const int numKeys = IARG(a, 2);
    GB_SCRATCH_OR_RETURN(S, a[3], n, numKeys);
    const long long NB = S.buckets;
    for (long long c = 0; c < S.chunks; ++c) {
        unsigned *h = S.hist + c * NB;
        memset(h, 0, sizeof(unsigned) * (size_t)NB);
        unsigned err = 0;
        for (long long r = c * PTK_GB_CHUNK; r < (c + 1) * PTK_GB_CHUNK && r < n; ++r) {
            if (keys[r] < 0 || keys[r] > numKeys) ++err;
            else ++h[keys[r] / PTK_GB_BUCKET];
        }
        S.berr[c] = err;
    }
}

/* GroupbyBucketScan(scratch, n, numKeys): per bucket, the rows of the earlier chunks (in place)
 * and the bucket's rows. */
void GroupbyBucketScan(unsigned int n_, void **a) {
    (void)n_;
    const long long n = IARG(a, 1);
    const int numKeys = IARG(a, 2);
    GB_SCRATCH_OR_RETURN(S, a[0], n, numKeys);
    const long long NB = S.buckets;
    for (long long b = 0; b < NB; ++b) {
        unsigned run = 0;
        for (long long c = 0; c < S.chunks; ++c) {
            const unsigned v = S.hist[c * NB + b];
            S.hist[c * NB + b] = run;
            run += v;
        }
        S.btot[b] = run;
    }
}

/* GroupbyPartition(keys, vals, n, numKeys, scratch): bucket and tile starts, the tile table, the
 * invalid keys; each valid row's low key byte and value at bstart[bucket] + its rank among the
 * bucket's rows (row order). */
void GroupbyPartition(unsigned int n_, void **a) {
    (void)n_;
    const int32_t *keys = (const int32_t *)a[0], *vals = (const int32_t *)a[1];
    const long long n = IARG(a, 2);
    const int numKeys = IARG(a, 3);
    GB_SCRATCH_OR_RETURN(S, a[4], n, numKeys);
    const long long NB = S.buckets;
    unsigned rows = 0, tiles = 0;
    for (long long b = 0; b < NB; ++b) {
        S.bstart[b] = rows;
        S.tstart[b] = tiles;
        const unsigned nt = (S.btot[b] + PTK_GB_TILE - 1) / PTK_GB_TILE;
        for (unsigned t = 0; t < nt; ++t) {
            const unsigned r0 = rows + t * PTK_GB_TILE, end = rows + S.btot[b];
            unsigned *ti = S.tinfo + 4ull * (tiles + t);
            ti[0] = (unsigned)b;
            ti[1] = r0;
            ti[2] = r0 + PTK_GB_TILE < end ? r0 + PTK_GB_TILE : end;
            ti[3] = 0;
        }
        rows += S.btot[b];
        tiles += nt;
    }
    S.bstart[NB] = rows;
    S.tstart[NB] = tiles;
    uint64_t err = 0;
    for (long long c = 0; c < S.chunks; ++c) err += S.berr[c];
    S.head[0] = err;
    unsigned *dest = (unsigned *)malloc(sizeof(unsigned) * (size_t)NB);
    for (long long c = 0; c < S.chunks; ++c) {
        for (long long b = 0; b < NB; ++b) dest[b] = S.bstart[b] + S.hist[c * NB + b];
        for (long long r = c * PTK_GB_CHUNK; r < (c + 1) * PTK_GB_CHUNK && r < n; ++r) {
            const int32_t k = keys[r];
            if (k < 0 || k > numKeys) continue;
            const unsigned g = dest[k / PTK_GB_BUCKET]++;
            S.tlow[g] = (unsigned char)(k & (PTK_GB_BUCKET - 1));
            S.tval[g] = vals[r];
        }
    }
    free(dest);
}

/* GroupbyKeyCount(scratch, n, numKeys): rows per key (low byte) per tile. */
void GroupbyKeyCount(unsigned int n_, void **a) {
    (void)n_;
    const long long n = IARG(a, 1);
    const int numKeys = IARG(a, 2);
    GB_SCRATCH_OR_RETURN(S, a[0], n, numKeys);
    for (unsigned t = 0; t < S.tstart[S.buckets]; ++t) {
        unsigned *c2 = S.c2 + (size_t)t * PTK_GB_BUCKET;
        memset(c2, 0, sizeof(unsigned) * PTK_GB_BUCKET);
        for (unsigned r = S.tinfo[4ull * t + 1]; r < S.tinfo[4ull * t + 2]; ++r) ++c2[S.tlow[r]];
    }
}

/* GroupbyKeyScan(scratch, n, numKeys, offsets): per key, the rows in the bucket's earlier tiles
 * (in place), its count, its first output row; the non-empty keys per half bucket. */
void GroupbyKeyScan(unsigned int n_, void **a) {
    (void)n_;
    const long long n = IARG(a, 1);
    const int numKeys = IARG(a, 2);
    int64_t *offsets = (int64_t *)a[3];
    GB_SCRATCH_OR_RETURN(S, a[0], n, numKeys);
    for (long long b = 0; b < S.buckets; ++b) {
        unsigned row = S.bstart[b], nz[2] = {0, 0};
        for (int k = 0; k < PTK_GB_BUCKET; ++k) {
            unsigned run = 0;
            for (unsigned t = S.tstart[b]; t < S.tstart[b + 1]; ++t) {
                unsigned *c = S.c2 + (size_t)t * PTK_GB_BUCKET + k;
                const unsigned v = *c;
                *c = run;
                run += v;
            }
            const long long key = b * PTK_GB_BUCKET + k;
            S.kcnt[key] = key <= numKeys ? run : 0u;
            if (key <= numKeys) {
                offsets[key] = row;
                row += run;
                nz[k / (PTK_GB_BUCKET / 2)] += run ? 1u : 0u;
            }
        }
        S.nz[2 * b] = nz[0];
        S.nz[2 * b + 1] = nz[1];
    }
}

/* GroupbyReshuffle(scratch, n, numKeys, offsets, outKeys, outVals): each tile's rows at
 * offsets[key] + the key's rows in the earlier tiles + its rank in the tile (row order); the
 * tile's per-key sums. */
void GroupbyReshuffle(unsigned int n_, void **a) {
    (void)n_;
    const long long n = IARG(a, 1);
    const int numKeys = IARG(a, 2);
    const int64_t *offsets = (const int64_t *)a[3];
    int32_t *outKeys = (int32_t *)a[4], *outVals = (int32_t *)a[5];
    GB_SCRATCH_OR_RETURN(S, a[0], n, numKeys);
    int64_t cur[PTK_GB_BUCKET];
    for (unsigned t = 0; t < S.tstart[S.buckets]; ++t) {
        const long long b = S.tinfo[4ull * t], key0 = b * PTK_GB_BUCKET;
        uint64_t *s2 = S.s2 + (size_t)t * PTK_GB_BUCKET;
        for (int k = 0; k < PTK_GB_BUCKET; ++k) {
            cur[k] = (key0 + k <= numKeys ? offsets[key0 + k] : 0) + S.c2[(size_t)t * PTK_GB_BUCKET + k];
            s2[k] = 0;
        }
        for (unsigned r = S.tinfo[4ull * t + 1]; r < S.tinfo[4ull * t + 2]; ++r) {
            const unsigned k = S.tlow[r];
            const int64_t g = cur[k]++;
            outKeys[g] = (int32_t)(key0 + k);
            outVals[g] = S.tval[r];
            s2[k] += (uint64_t)(int64_t)S.tval[r];
        }
    }
}

/* GroupbyTotals(scratch, n, numKeys, table, groups): per key its count and its sum over the
 * bucket's tiles (tile order); the group records in key order and the header. */
void GroupbyTotals(unsigned int n_, void **a) {
    (void)n_;
    const long long n = IARG(a, 1);
    const int numKeys = IARG(a, 2);
    int64_t *table = (int64_t *)a[3], *groups = (int64_t *)a[4];
    GB_SCRATCH_OR_RETURN(S, a[0], n, numKeys);
    uint64_t *cnt = (uint64_t *)table, *sum = cnt + (numKeys + 1), *err = cnt + 3 * ((size_t)numKeys + 1);
    uint64_t G = 0;
    for (long long b = 0; b < S.buckets; ++b)
        for (int k = 0; k < PTK_GB_BUCKET; ++k) {
            const long long key = b * PTK_GB_BUCKET + k;
            if (key > numKeys) break;
            uint64_t v = 0;
            for (unsigned t = S.tstart[b]; t < S.tstart[b + 1]; ++t) v += S.s2[(size_t)t * PTK_GB_BUCKET + k];
            cnt[key] = S.kcnt[key];
            sum[key] = v;
            if (S.kcnt[key]) {
                int64_t *e = groups + 3 + 3 * G;
                e[0] = key;
                e[1] = (int64_t)S.kcnt[key];
                e[2] = (int64_t)v;
                ++G;
            }
        }
    *err = S.head[0];
    groups[0] = (int64_t)G;
    groups[1] = (int64_t)S.bstart[S.buckets];
    groups[2] = (int64_t)S.head[0];
}

/* CoalesceGroups (HostTasks/HostTasks.cpp:131-197) with the ABI of its coalesce_int_float entry
 * (:164-197): (float* values, int n, int* groupIDs, int numGroups, int* counts, void* out,
 * int* meta).  The values arrive grouped; group i is written as its key (groupIDs at the group's
 * first element) followed by its counts[i] values, groups one after another, and meta[i] =
 * counts[i].  The reference's loop never advances its output cursor past a group (every group
 * is written at the start of the output); this restatement writes the layout the function
 * documents, groups consecutive. */
void coalesce_int_float(unsigned int nArgs, void **a) {
    (void)nArgs;
    const float *in = (const float *)a[0];
    const int n = IARG(a, 1);
    const int32_t *ids = (const int32_t *)a[2];
    const int groups = IARG(a, 3);
    const int32_t *counts = (const int32_t *)a[4];
    unsigned char *out = (unsigned char *)a[5];
    int32_t *meta = (int32_t *)a[6];
    int pos = 0;
    for (int g = 0; g < groups && pos < n; ++g) {
        const int32_t c = counts[g];
        memcpy(out, &ids[pos], sizeof(int32_t));
        out += sizeof(int32_t);
        for (int j = 0; j < c && pos < n; ++j, ++pos) {
            memcpy(out, &in[pos], sizeof(float));
            out += sizeof(float);
        }
        meta[g] = c;
    }
}

/* GroupbyCoalesce: the same layout over this pipeline's outputs, with the HIP kernel's ABI
 * (int* outVals, int* outKeys, long long* groups, long long* offsets, int* out, int* meta):
 * G and the valid rows come from the groups header, the counts from its records. */
void GroupbyCoalesce(unsigned int nArgs, void **a) {
    (void)nArgs;
    const int64_t *groups = (const int64_t *)a[2];
    const int G = (int)groups[0], rows = (int)groups[1];
    int32_t *counts = (int32_t *)malloc(sizeof(int32_t) * (size_t)(G > 0 ? G : 1));
    for (int g = 0; g < G; ++g) counts[g] = (int32_t)groups[3 + 3 * (size_t)g + 1];
    void *args[7] = {a[0], (void *)(intptr_t)rows, a[1], (void *)(intptr_t)G, counts, a[4], a[5]};
    coalesce_int_float(7, args);
    free(counts);
}

/* GroupbyCoalesceSizes(groups, &coalescedBytes, &groupCountsBytes): the exact byte sizes of the
 * coalesced layout (4 (rows + G)) and of the per-group counts (4 G), at least 4 each. */
void GroupbyCoalesceSizes(unsigned int n, void **a) {
    (void)n;
    const long long *groups = (const long long *)a[0];
    const long long G = groups[0], rows = groups[1];
    *(int *)a[1] = (int)(4 * (rows + G > 0 ? rows + G : 1));
    *(int *)a[2] = (int)(4 * (G > 0 ? G : 1));
}
