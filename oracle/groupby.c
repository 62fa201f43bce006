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

/* The chunk / bucket pipeline of the HIP kernels (include/ptk_groupby.h) restated on the host:
 * the same scratch layout, with each chunk's distinct keys in ascending key order (the device
 * writes them in any order inside a bucket; nothing downstream depends on it); merge bases,
 * group bases and ranks follow row order, as on the device. */
static int cmp_i32(const void *x, const void *y) {
    const int32_t a = *(const int32_t *)x, b = *(const int32_t *)y;
    return (a > b) - (a < b);
}

void GroupbyAccumulate(unsigned int n_, void **a) {
    (void)n_;
    const int32_t *keys = (const int32_t *)a[0], *vals = (const int32_t *)a[1];
    const long long n = IARG(a, 2);
    const int numKeys = IARG(a, 3);
    if (n <= 0 || numKeys < 0 || (long long)numKeys + 1 > PTK_GB_MAX_KEYS) return;
    PTK_GB_SCRATCH S = ptk_gb_scratch(a[4], n, numKeys);
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * PTK_GB_CHUNK);
    for (long long c = 0; c < S.chunks; ++c) {
        const long long r0 = c * PTK_GB_CHUNK, r1 = r0 + PTK_GB_CHUNK < n ? r0 + PTK_GB_CHUNK : n;
        unsigned err = 0, m = 0;
        for (long long r = r0; r < r1; ++r) {
            if (keys[r] < 0 || keys[r] > numKeys) ++err;
            else tmp[m++] = keys[r];
        }
        qsort(tmp, m, sizeof(int32_t), cmp_i32);
        PTK_GB_ENTRY *e = S.entry + c * PTK_GB_CHUNK;
        unsigned d = 0;
        for (unsigned i = 0; i < m; ++i)
            if (i == 0 || tmp[i] != tmp[i - 1]) {
                e[d].key = tmp[i];
                e[d].count = 0;
                e[d].sum = 0;
                ++d;
            }
        for (long long r = r0; r < r1; ++r) {  /* add each row to its entry (binary search) */
            const int32_t k = keys[r];
            if (k < 0 || k > numKeys) continue;
            unsigned lo = 0, hi = d;
            while (hi - lo > 1) {
                const unsigned mid = (lo + hi) / 2;
                if (e[mid].key <= k) lo = mid;
                else hi = mid;
            }
            e[lo].count += 1;
            e[lo].sum = (long long)((uint64_t)e[lo].sum + (uint64_t)(int64_t)vals[r]);
        }
        unsigned *boff = S.boff + c * (S.buckets + 1);
        unsigned i = 0;
        for (long long b = 0; b <= S.buckets; ++b) {  /* entries of buckets < b */
            while (i < d && e[i].key / PTK_GB_BUCKET < b) ++i;
            boff[b] = i;
        }
        S.berr[c] = err;
    }
    free(tmp);
}

/* GroupbyMerge(scratch, n, numKeys): per merge group, each entry's base (the key's rows in the
 * group's earlier chunks) replaces its sum; the group's per-key totals go to gcnt / gsum. */
void GroupbyMerge(unsigned int n_, void **a) {
    (void)n_;
    const long long n = IARG(a, 1);
    const int numKeys = IARG(a, 2);
    if (n < 0 || numKeys < 0 || (long long)numKeys + 1 > PTK_GB_MAX_KEYS) return;
    PTK_GB_SCRATCH S = ptk_gb_scratch(a[0], n, numKeys);
    const long long KP = S.buckets * PTK_GB_BUCKET;
    memset(S.gcnt, 0, sizeof(unsigned) * (size_t)(S.groups * KP));
    memset(S.gsum, 0, sizeof(uint64_t) * (size_t)(S.groups * KP));
    for (long long c = 0; c < S.chunks; ++c) {  /* within a merge group, chunks in row order */
        unsigned *gc = S.gcnt + (c / PTK_GB_GROUP) * KP;
        unsigned long long *gs = S.gsum + (c / PTK_GB_GROUP) * KP;
        const unsigned d = S.boff[c * (S.buckets + 1) + S.buckets];
        for (unsigned i = 0; i < d; ++i) {
            PTK_GB_ENTRY *e = S.entry + c * PTK_GB_CHUNK + i;
            gs[e->key] += (uint64_t)e->sum;
            e->sum = (long long)gc[e->key];  /* the entry's base inside its group replaces its sum */
            gc[e->key] += e->count;
        }
    }
}

/* GroupbyTotals(scratch, n, numKeys, table): groups in row order give each key's per-group base
 * (replacing gcnt) and its count and sum; the bucket totals and the error count. */
void GroupbyTotals(unsigned int n_, void **a) {
    (void)n_;
    const long long n = IARG(a, 1);
    const int numKeys = IARG(a, 2);
    int64_t *table = (int64_t *)a[3];
    if (n < 0 || numKeys < 0 || (long long)numKeys + 1 > PTK_GB_MAX_KEYS) return;
    PTK_GB_SCRATCH S = ptk_gb_scratch(a[0], n, numKeys);
    const long long KP = S.buckets * PTK_GB_BUCKET;
    uint64_t *cnt = (uint64_t *)table, *sum = cnt + (numKeys + 1), *err = cnt + 3 * ((size_t)numKeys + 1);
    for (long long k = 0; k < KP; ++k) {
        uint64_t t = 0, v = 0;
        for (long long g = 0; g < S.groups; ++g) {
            const unsigned c = S.gcnt[g * KP + k];
            S.gcnt[g * KP + k] = (unsigned)t;
            t += c;
            v += S.gsum[g * KP + k];
        }
        if (k <= numKeys) {
            cnt[k] = t;
            sum[k] = v;
        }
    }
    uint64_t errors = 0;
    for (long long c = 0; c < S.chunks; ++c) errors += S.berr[c];
    *err = errors;
    for (long long b = 0; b < S.buckets; ++b) {
        uint64_t rows = 0, nz = 0;
        for (long long k = b * PTK_GB_BUCKET; k < (b + 1) * PTK_GB_BUCKET && k <= numKeys; ++k) {
            rows += cnt[k];
            nz += cnt[k] ? 1 : 0;
        }
        S.btot[2 * b] = (int64_t)rows;
        S.btot[2 * b + 1] = (int64_t)nz;
    }
}

void GroupbyOffsets(unsigned int n, void **a) {
    (void)n;
    gb_offsets((const int64_t *)a[0], IARG(a, 1), (int64_t *)a[3], (int64_t *)a[4]);
}

void GroupbyReshuffle(unsigned int n_, void **a) {
    (void)n_;
    const int32_t *keys = (const int32_t *)a[0], *vals = (const int32_t *)a[1];
    const long long n = IARG(a, 2);
    const int numKeys = IARG(a, 3);
    const int64_t *offsets = (const int64_t *)a[4];
    int32_t *outKeys = (int32_t *)a[6], *outVals = (int32_t *)a[7];
    if (n <= 0 || numKeys < 0 || (long long)numKeys + 1 > PTK_GB_MAX_KEYS) return;
    PTK_GB_SCRATCH S = ptk_gb_scratch(a[5], n, numKeys);
    int64_t *next = (int64_t *)malloc(sizeof(int64_t) * PTK_GB_CHUNK);
    for (long long c = 0; c < S.chunks; ++c) {
        const long long r0 = c * PTK_GB_CHUNK, r1 = r0 + PTK_GB_CHUNK < n ? r0 + PTK_GB_CHUNK : n;
        const PTK_GB_ENTRY *e = S.entry + c * PTK_GB_CHUNK;
        const unsigned d = S.boff[c * (S.buckets + 1) + S.buckets];
        const unsigned *gbase = S.gcnt + (c / PTK_GB_GROUP) * (S.buckets * PTK_GB_BUCKET);
        for (unsigned i = 0; i < d; ++i) next[i] = offsets[e[i].key] + gbase[e[i].key] + e[i].sum;  /* sum: the merge base */
        for (long long r = r0; r < r1; ++r) {
            const int32_t k = keys[r];
            if (k < 0 || k > numKeys || d == 0) continue;
            unsigned lo = 0, hi = d;
            while (hi - lo > 1) {
                const unsigned mid = (lo + hi) / 2;
                if (e[mid].key <= k) lo = mid;
                else hi = mid;
            }
            const int64_t pos = next[lo]++;
            outKeys[pos] = k;
            outVals[pos] = vals[r];
        }
    }
    free(next);
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
