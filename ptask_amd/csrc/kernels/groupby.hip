// groupby.hip -- the groupby graph's task kernels (code object groupby.hsaco).
//
// The reference graph (ptaskUnitTest/graphgroupby.cpp:106-387) counts unique keys, builds
// per-key counts in a device hash table, prefix-sums them (single block) and reshuffles the
// input so equal keys are contiguous; its kernels are not in the tree.  This pipeline keeps that
// shape, adds an int32 value column aggregated to int64 sum and count per key, and replaces the
// device hash table by per-chunk LDS tables merged per key bucket (no global atomics on the data;
// every output, the order of rows inside a group included, is bit-reproducible; the bodies are
// in groupby_body.h, the layouts in include/ptk_groupby.h):
//
//   GroupbyAccumulate(int* keys, int* vals, int n, int numKeys, void* scratch)     -> scratch
//   GroupbyMerge(void* scratch, int n, int numKeys)                      scratch in/out
//   GroupbyTotals(void* scratch, int n, int numKeys, long long* table) -> table; scratch in/out
//   GroupbyOffsets(long long* table, int numKeys, void* scratch, long long* offsets,
//                  long long* groups)
//   GroupbyReshuffle(int* keys, int* vals, int n, int numKeys, long long* offsets,
//                    void* scratch, int* outKeys, int* outVals)
//
// Launch: Accumulate takes any grid of 64k-thread blocks (blocks stride over the
// ceil(n / PTK_GB_CHUNK) chunks); Reshuffle any grid of PTK_GB_PLACE_THREADS-thread blocks
// (the same chunks); Merge any grid of PTK_GB_BUCKET-thread blocks (blocks stride over the
// buckets x merge groups units); Totals and Offsets blocks of exactly PTK_GB_BUCKET threads, any
// grid.  Keys lie in [0, numKeys] as in the
// reference test (graphgroupby.cpp:428); a key outside is counted in the table's error slot and
// otherwise ignored, so no access leaves its array.  numKeys + 1 <= PTK_GB_MAX_KEYS (checked).
#include "groupby_body.h"

using namespace ptk_groupby;

extern "C" __global__ void __launch_bounds__(1024)
GroupbyAccumulate(const int *__restrict__ keys, const int *__restrict__ vals, int n, int numKeys, void *scratch) {
    __shared__ GbChunkLds L;
    if (!gb_keys_ok(numKeys) || n <= 0 || (blockDim.x & 63) != 0) return;
    const PTK_GB_SCRATCH S = ptk_gb_scratch(scratch, n, numKeys);
    for (long long c = blockIdx.x; c < S.chunks; c += gridDim.x) gb_accumulate_chunk(keys, vals, n, numKeys, c, S, L);
}

// Units (bucket, merge group), bucket-major; neighbouring units share cache lines of the entry
// and offset arrays, so each XCD (blocks are dealt to the 8 XCDs round-robin) takes a contiguous
// run of units.  Blocks of exactly PTK_GB_BUCKET threads, any grid.
extern "C" __global__ void __launch_bounds__(PTK_GB_BUCKET) GroupbyMerge(void *scratch, int n, int numKeys) {
    __shared__ GbMergeLds L;
    if (!gb_keys_ok(numKeys) || n < 0 || blockDim.x != PTK_GB_BUCKET) return;
    const PTK_GB_SCRATCH S = ptk_gb_scratch(scratch, n, numKeys);
    const long long units = S.buckets * S.groups;
    for (long long u = gb_xcd_block(blockIdx.x, gridDim.x); u < units; u += gridDim.x) gb_merge_unit(S, u, L);
}

// One block per bucket (any grid; blocks stride over the buckets), blocks of exactly
// PTK_GB_BUCKET threads.
extern "C" __global__ void __launch_bounds__(PTK_GB_BUCKET) GroupbyTotals(void *scratch, int n, int numKeys, long long *table) {
    __shared__ unsigned long long red[16];
    if (!gb_keys_ok(numKeys) || n < 0 || blockDim.x != PTK_GB_BUCKET) return;
    const PTK_GB_SCRATCH S = ptk_gb_scratch(scratch, n, numKeys);
    for (long long r = blockIdx.x; r < S.buckets; r += gridDim.x) gb_totals_bucket(S, numKeys, (int)r, table, red);
}

extern "C" __global__ void __launch_bounds__(PTK_GB_BUCKET)
GroupbyOffsets(long long *table, int numKeys, void *scratch, long long *offsets, long long *groups) {
    __shared__ GbOffsetsLds L;
    if (!gb_keys_ok(numKeys) || blockDim.x != PTK_GB_BUCKET) return;
    // the scratch prefix holding the bucket totals does not depend on n
    const PTK_GB_SCRATCH S = ptk_gb_scratch(scratch, 0, numKeys);
    for (long long r = blockIdx.x; r < S.buckets; r += gridDim.x) gb_offsets_bucket(table, numKeys, S, (int)r, offsets, groups, L);
}

extern "C" __global__ void __launch_bounds__(PTK_GB_PLACE_THREADS)
GroupbyReshuffle(const int *__restrict__ keys, const int *__restrict__ vals, int n, int numKeys,
                 const long long *__restrict__ offsets, void *scratch, int *__restrict__ outKeys,
                 int *__restrict__ outVals) {
    __shared__ GbPlaceLds L;
    if (!gb_keys_ok(numKeys) || n <= 0 || blockDim.x != PTK_GB_PLACE_THREADS) return;
    const PTK_GB_SCRATCH S = ptk_gb_scratch(scratch, n, numKeys);
    // Consecutive chunks place a key's rows in consecutive slots of its group, so each XCD takes
    // a contiguous run of chunks: the partial lines of rare keys' groups then fill up in one L2
    // before they are written back.
    const unsigned G = gridDim.x < S.chunks ? gridDim.x : (unsigned)S.chunks;
    if (blockIdx.x >= G) return;
    for (long long c = gb_xcd_block(blockIdx.x, G); c < S.chunks; c += G)
        gb_place_chunk(keys, vals, n, numKeys, c, offsets, S, outKeys, outVals, L);
}

// The reference's CoalesceGroups host task (HostTasks/HostTasks.cpp:131-197) over this
// pipeline's outputs: primary output, per group in key order, the key followed by the group's
// values (rows + G words); meta output, the count of each group.  Values are copied as 4-byte
// words (int32 here; the reference instantiates float).  Any grid of 64k-thread blocks; blocks
// stride over chunks of PTK_GB_CHUNK grouped rows.
//   GroupbyCoalesce(int* outVals, int* outKeys, long long* groups, long long* offsets,
//                   int* coalesced, int* groupCounts)
extern "C" __global__ void __launch_bounds__(1024)
GroupbyCoalesce(const int *__restrict__ vals, const int *__restrict__ keys, const long long *__restrict__ groups,
                const long long *__restrict__ offsets, int *__restrict__ out, int *__restrict__ meta) {
    __shared__ GbCoalesceLds L;
    if ((blockDim.x & 63) != 0) return;
    const long long chunks = (groups[1] + PTK_GB_CHUNK - 1) / PTK_GB_CHUNK;
    for (long long c = blockIdx.x; c < chunks; c += gridDim.x) gb_coalesce_chunk(vals, keys, groups, offsets, c, out, meta, L);
}

// Byte sizes of GroupbyCoalesce's two outputs for a graph that allocates them exactly (through
// MF_ALLOCATION_SIZE meta ports, MetaPort.cpp:724): coalesced = 4 (rows + G), groupCounts = 4 G,
// from the groups header {G, rows, errors}; at least 4 bytes each.  One thread.
//   GroupbyCoalesceSizes(long long* groups, int* coalescedBytes, int* groupCountsBytes)
extern "C" __global__ void __launch_bounds__(64)
GroupbyCoalesceSizes(const long long *__restrict__ groups, int *coalescedBytes, int *groupCountsBytes) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const long long G = groups[0], rows = groups[1];
    *coalescedBytes = (int)(4 * (rows + G > 0 ? rows + G : 1));
    *groupCountsBytes = (int)(4 * (G > 0 ? G : 1));
}
