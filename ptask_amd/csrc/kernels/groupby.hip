// groupby.hip -- the groupby graph's task kernels (code object groupby.hsaco).
//
// The reference graph (ptaskUnitTest/graphgroupby.cpp:106-387) counts unique keys, builds
// per-key counts in a device hash table, prefix-sums them (single block) and reshuffles the
// input so equal keys are contiguous; its kernels are not in the tree.  This pipeline keeps that
// shape and adds an int32 value column aggregated to int64 sum and count per key:
//
//   GroupbyClear(long long* table, int numKeys)                       -> table zeroed
//   GroupbyAccumulate(int* keys, int* vals, int n, int numKeys, long long* table)  [table in/out]
//   GroupbyOffsets(long long* table, int numKeys, long long* offsets, long long* groups)
//   GroupbyReshuffle(int* keys, int* vals, int n, int numKeys, long long* offsets,
//                    long long* table, int* outKeys, int* outVals)     [table in/out: cursors]
//
// Keys lie in [0, numKeys] as in the reference test (graphgroupby.cpp:428); a key outside is
// counted in the table's error slot and otherwise ignored, so no access leaves its array.
// table layout (int64): count[numKeys+1] | sum[numKeys+1] | cursor[numKeys+1] | errors[1].
// groups layout (int64): header {G, rows, errors} then G triples {key, count, sum} in
// ascending key order.  offsets[k] = rows with a smaller key (exclusive scan of count).
//
// Sums and counts use exact int64 (wrapping) additions, whose result does not depend on the
// order of the atomics: every output except the order of values inside one group is
// bit-reproducible.
#include "groupby_body.h"

using namespace ptk_groupby;

extern "C" __global__ void __launch_bounds__(1024) GroupbyClear(long long *table, int numKeys) {
    if (numKeys < 0) return;
    gb_clear(table, numKeys, (long long)blockIdx.x * blockDim.x + threadIdx.x, (long long)gridDim.x * blockDim.x);
}

// Each block aggregates a contiguous chunk of the rows (see gb_accumulate_rows).
extern "C" __global__ void __launch_bounds__(1024)
GroupbyAccumulate(const int *__restrict__ keys, const int *__restrict__ vals, int n, int numKeys,
                  long long *table) {
    if (numKeys < 0 || n <= 0) return;
    const long long per = ((long long)n + gridDim.x - 1) / gridDim.x;
    const long long r0 = per * blockIdx.x, r1 = r0 + per < n ? r0 + per : n;
    gb_accumulate_rows(keys, vals, r0, r1, numKeys, table);
}

// ---- one-block scans -------------------------------------------------------------------------
// One block: offsets and the group list (see gb_offsets_block).
extern "C" __global__ void __launch_bounds__(1024)
GroupbyOffsets(long long *table, int numKeys, long long *offsets, long long *groups) {
    if ((blockDim.x & 63) != 0 || blockIdx.x != 0 || numKeys < 0) return;
    gb_offsets_block(table, numKeys, offsets, groups);
}

// Each block reshuffles a contiguous chunk of the rows (see gb_reshuffle_rows).
extern "C" __global__ void __launch_bounds__(1024)
GroupbyReshuffle(const int *__restrict__ keys, const int *__restrict__ vals, int n, int numKeys,
                 const long long *__restrict__ offsets, long long *table, int *__restrict__ outKeys,
                 int *__restrict__ outVals) {
    if (numKeys < 0 || n <= 0) return;
    const long long per = ((long long)n + gridDim.x - 1) / gridDim.x;
    const long long r0 = per * blockIdx.x, r1 = r0 + per < n ? r0 + per : n;
    gb_reshuffle_rows(keys, vals, n, r0, r1, numKeys, offsets, table, outKeys, outVals);
}
