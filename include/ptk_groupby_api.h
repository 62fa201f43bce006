/*
 * ptk_groupby_api.h -- the groupby pipeline as direct library calls (libptk_groupby.so), for
 * callers that hold their rows in device memory outside a PTask graph.
 *
 * One call runs the same seven task kernels the groupby graph dispatches (GroupbyHistogram,
 * GroupbyBucketScan, GroupbyPartition, GroupbyKeyCount, GroupbyKeyScan, GroupbyReshuffle,
 * GroupbyTotals; layouts in ptk_groupby.h) on the caller's stream, in stream order, without synchronising.  Results are bitwise those of the
 * graph: per-key int64 counts and sums, the groups block {G, rows, errors, then G {key, count,
 * sum} in key order}, offsets, and the rows grouped by ascending key with the rows of a group in
 * input order.  Keys lie in [0, numKeys]; other keys are counted as errors and dropped.
 *
 * All pointers are device pointers; stream is a hipStream_t (NULL: the null stream).  Every
 * function returns 0 on success or a PTK_GB_E* code and never throws; ptk_groupby_last_error
 * gives the text of the calling thread's last failure.
 */
#ifndef PTK_GROUPBY_API_H
#define PTK_GROUPBY_API_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PTK_GB_OK = 0,
    PTK_GB_EINVAL = 1,   /* bad pointer, row count or key range */
    PTK_GB_ESPACE = 2,   /* the workspace (or the static no-allocation workspace) is too small */
    PTK_GB_EHIP = 3,     /* a HIP call failed */
};

/* Bytes of device workspace ptk_groupby_run needs for n rows over keys [0, numKeys] (the
 * per-key table and the chunk / tile scratch of ptk_groupby.h); -1 for a bad shape. */
long long ptk_groupby_workspace_bytes(long long n, int numKeys);

/* The groupby graph (ptaskUnitTest/graphgroupby.cpp:332-387: count unique keys, allocate,
 * count, prefix-sum, reshuffle; its output invariant :394-419) as one call, plus the per-key
 * int64 count and sum of an int32 value column.
 *   keys, vals  n int32 rows (1 <= n <= 2^31 - 1)
 *   workspace   at least ptk_groupby_workspace_bytes(n, numKeys) bytes, 16-byte aligned; its
 *               contents on entry do not matter
 *   offsets     numKeys + 1 int64: rows with a smaller key
 *   groups      3 + 3 (numKeys + 1) int64 (only the header and G records are written)
 *   outKeys, outVals  n int32 each (the first rows - errors are written) */
int ptk_groupby_run(const int *keys, const int *vals, long long n, int numKeys, void *workspace,
                    unsigned long long workspaceBytes, long long *offsets, long long *groups, int *outKeys,
                    int *outVals, void *stream);

/* No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  As ptk_groupby_run,
 * but the workspace is a static device array of PTK_GB_NOALLOC_BYTES inside the library, one per
 * GPU, held until the process ends; the call allocates nothing and fails with PTK_GB_ESPACE when
 * ptk_groupby_workspace_bytes(n, numKeys) exceeds it (about 11M rows over 65536 keys).  The
 * array keeps the last call's contents; nothing in it needs clearing between calls.  Calls that
 * overlap on one GPU (other streams, other threads) share the array and race: serialise them. */
#define PTK_GB_NOALLOC_BYTES (64ull << 20)
unsigned long long ptk_groupby_noalloc_capacity(void);
int ptk_groupby_run_noalloc(const int *keys, const int *vals, long long n, int numKeys, long long *offsets,
                            long long *groups, int *outKeys, int *outVals, void *stream);

const char *ptk_groupby_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* PTK_GROUPBY_API_H */
