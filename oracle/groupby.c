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

void GroupbyClear(unsigned int n, void **a) {
    (void)n;
    gb_clear((int64_t *)a[0], IARG(a, 1));
}

void GroupbyAccumulate(unsigned int n, void **a) {
    (void)n;
    gb_accumulate((const int32_t *)a[0], (const int32_t *)a[1], IARG(a, 2), IARG(a, 3), (int64_t *)a[4]);
}

void GroupbyOffsets(unsigned int n, void **a) {
    (void)n;
    gb_offsets((const int64_t *)a[0], IARG(a, 1), (int64_t *)a[2], (int64_t *)a[3]);
}

void GroupbyReshuffle(unsigned int n, void **a) {
    (void)n;
    gb_reshuffle((const int32_t *)a[0], (const int32_t *)a[1], IARG(a, 2), IARG(a, 3),
                 (const int64_t *)a[4], (int64_t *)a[5], (int32_t *)a[6], (int32_t *)a[7]);
}
