// groupby.hip -- the groupby graph's task kernels (code object groupby.hsaco).
//
// The reference graph (ptaskUnitTest/graphgroupby.cpp:106-387) counts unique keys, builds
// per-key counts in a device hash table, prefix-sums them (single block) and reshuffles the
// input so equal keys are contiguous; its kernels are not in the tree.  This pipeline keeps that
// shape (count, prefix, reshuffle), adds an int32 value column aggregated to int64 sum and count
// per key, and orders the rows by a stable two-level radix partition (buckets of 256 keys, then
// keys) with no global atomics: every output, the order of rows inside a group included, is
// bit-reproducible.  The bodies are in groupby_body.h, the layouts in include/ptk_groupby.h.
//
//   GroupbyHistogram(int* keys, int n, int numKeys, void* scratch)              -> scratch
//   GroupbyBucketScan(void* scratch, int n, int numKeys)                         scratch in/out
//   GroupbyPartition(int* keys, int* vals, int n, int numKeys, void* scratch)    scratch in/out
//   GroupbyKeyCount(void* scratch, int n, int numKeys)                           scratch in/out
//   GroupbyKeyScan(void* scratch, int n, int numKeys, long long* offsets)        scratch in/out
//   GroupbyReshuffle(void* scratch, int n, int numKeys, long long* offsets, int* outKeys,
//                    int* outVals)                                               scratch in/out
//   GroupbyTotals(void* scratch, int n, int numKeys, long long* table, long long* groups)
//
// Launch: any grid for each (blocks stride over their units: chunks, buckets or tiles); blocks
// of exactly PTK_GB_HIST_THREADS (Histogram), PTK_GB_PART_THREADS (Partition),
// PTK_GB_COUNT_THREADS (KeyCount), PTK_GB_TILE_THREADS (Reshuffle), PTK_GB_SCAN_THREADS
// (BucketScan), PTK_GB_KEYSCAN_THREADS (KeyScan) and PTK_GB_TOTALS_THREADS (Totals) threads; a launch with another block size does nothing.  Keys lie in [0, numKeys] as in
// the reference test (graphgroupby.cpp:428); a key outside is counted in the table's error slot
// and otherwise ignored, so no access leaves its array.  numKeys + 1 <= PTK_GB_MAX_KEYS (checked).
#include "groupby_body.h"

using namespace ptk_groupby;

extern "C" __global__ void __launch_bounds__(PTK_GB_HIST_THREADS)
GroupbyHistogram(const int *__restrict__ keys, int n, int numKeys, void *scratch) {
    __shared__ GbHistLds L;
    if (!gb_keys_ok(numKeys) || n < 0 || blockDim.x != PTK_GB_HIST_THREADS) return;
    const PTK_GB_SCRATCH S = ptk_gb_scratch(scratch, n, numKeys);
    for (long long c = blockIdx.x; c < S.chunks; c += gridDim.x) gb_histogram_chunk(keys, n, numKeys, c, S, L);
}

extern "C" __global__ void __launch_bounds__(PTK_GB_SCAN_THREADS) GroupbyBucketScan(void *scratch, int n, int numKeys) {
    __shared__ unsigned red[16];
    if (!gb_keys_ok(numKeys) || n < 0 || blockDim.x != PTK_GB_SCAN_THREADS) return;
    const PTK_GB_SCRATCH S = ptk_gb_scratch(scratch, n, numKeys);
    for (long long b = blockIdx.x; b < S.buckets; b += gridDim.x) gb_bucketscan(S, (int)b, red);
}

// Consecutive chunks write neighbouring runs of each bucket, so each XCD (blocks are dealt to the
// 8 XCDs round-robin) takes a contiguous run of chunks: the partial lines at the runs' ends then
// fill up in one L2 before they are written back.
extern "C" __global__ void __launch_bounds__(PTK_GB_PART_THREADS, 4)
GroupbyPartition(const int *__restrict__ keys, const int *__restrict__ vals, int n, int numKeys, void *scratch) {
    __shared__ GbPartLds L;
    if (!gb_keys_ok(numKeys) || n < 0 || blockDim.x != PTK_GB_PART_THREADS) return;
    const PTK_GB_SCRATCH S = ptk_gb_scratch(scratch, n, numKeys);
    if (blockIdx.x == 0) gb_partition_starts(S, L);
    const unsigned G = gridDim.x < S.chunks ? gridDim.x : (unsigned)S.chunks;
    if (blockIdx.x >= G) return;
    gb_partition_chunks(keys, vals, n, numKeys, gb_xcd_block(blockIdx.x, G), S.chunks, G, S, L);
}

extern "C" __global__ void __launch_bounds__(PTK_GB_COUNT_THREADS) GroupbyKeyCount(void *scratch, int n, int numKeys) {
    __shared__ GbKeyCountLds L;
    if (!gb_keys_ok(numKeys) || n < 0 || blockDim.x != PTK_GB_COUNT_THREADS) return;
    const PTK_GB_SCRATCH S = ptk_gb_scratch(scratch, n, numKeys);
    const unsigned NT = S.tstart[S.buckets];
    for (unsigned t = blockIdx.x; t < NT; t += gridDim.x) gb_keycount_tile(S, t, L);
}

// Units of half a bucket (two per bucket); any grid.
extern "C" __global__ void __launch_bounds__(PTK_GB_KEYSCAN_THREADS)
GroupbyKeyScan(void *scratch, int n, int numKeys, long long *offsets) {
    __shared__ GbKeyScanLds L;
    if (!gb_keys_ok(numKeys) || n < 0 || blockDim.x != PTK_GB_KEYSCAN_THREADS) return;
    const PTK_GB_SCRATCH S = ptk_gb_scratch(scratch, n, numKeys);
    for (long long u = blockIdx.x; u < 2 * S.buckets; u += gridDim.x) gb_keyscan_half(S, numKeys, (int)(u / 2), (int)(u % 2), offsets, L);
}

// Consecutive tiles of a bucket place rows of the same keys next to each other, so each XCD takes
// a contiguous run of tiles (as in GroupbyPartition).
extern "C" __global__ void __launch_bounds__(PTK_GB_TILE_THREADS)
GroupbyReshuffle(void *scratch, int n, int numKeys, const long long *__restrict__ offsets, int *__restrict__ outKeys,
                 int *__restrict__ outVals) {
    __shared__ GbShuffleLds L;
    if (!gb_keys_ok(numKeys) || n < 0 || blockDim.x != PTK_GB_TILE_THREADS) return;
    const PTK_GB_SCRATCH S = ptk_gb_scratch(scratch, n, numKeys);
    const unsigned NT = S.tstart[S.buckets];
    const unsigned G = gridDim.x < NT ? gridDim.x : NT;
    if (blockIdx.x >= G) return;
    gb_reshuffle_tiles(S, numKeys, gb_xcd_block(blockIdx.x, G), NT, G, offsets, outKeys, outVals, n, L);
}

// Units of a quarter bucket (four per bucket); any grid.
extern "C" __global__ void __launch_bounds__(PTK_GB_TOTALS_THREADS)
GroupbyTotals(void *scratch, int n, int numKeys, long long *table, long long *groups) {
    __shared__ GbTotalsLds L;
    if (!gb_keys_ok(numKeys) || n < 0 || blockDim.x != PTK_GB_TOTALS_THREADS) return;
    const PTK_GB_SCRATCH S = ptk_gb_scratch(scratch, n, numKeys);
    for (long long u = blockIdx.x; u < 4 * S.buckets; u += gridDim.x) gb_totals_quarter(S, numKeys, (int)(u / 4), (int)(u % 4), table, groups, L);
}

// The reference's CoalesceGroups host task (HostTasks/HostTasks.cpp:131-197) over this
// pipeline's outputs: primary output, per group in key order, the key followed by the group's
// values (rows + G words); meta output, the count of each group.  Values are copied as 4-byte
// words (int32 here; the reference instantiates float).  Any grid of 64k-thread blocks; blocks
// stride over chunks of GB_COALESCE_ROWS grouped rows.
//   GroupbyCoalesce(int* outVals, int* outKeys, long long* groups, long long* offsets,
//                   int* coalesced, int* groupCounts)
extern "C" __global__ void __launch_bounds__(1024)
GroupbyCoalesce(const int *__restrict__ vals, const int *__restrict__ keys, const long long *__restrict__ groups,
                const long long *__restrict__ offsets, int *__restrict__ out, int *__restrict__ meta) {
    __shared__ GbCoalesceLds L;
    if ((blockDim.x & 63) != 0) return;
    const long long chunks = (groups[1] + GB_COALESCE_ROWS - 1) / GB_COALESCE_ROWS;
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
