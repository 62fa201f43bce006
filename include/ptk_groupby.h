/*
 * ptk_groupby.h -- layouts shared by the groupby task kernels, their host-task forms and the
 * applications that build the groupby graph.
 *
 * The reference graph (ptaskUnitTest/graphgroupby.cpp:106-387) counts keys in a device hash
 * table, prefix-sums the counts and reshuffles the input so that equal keys are contiguous; its
 * kernels are not in the tree.  This pipeline keeps that shape, adds an int32 value column
 * aggregated to an int64 count and sum per key, and uses no global atomics:
 *
 *   GroupbyAccumulate(int* keys, int* vals, int n, int numKeys, void* scratch)   -> scratch
 *   GroupbyMerge(void* scratch, int n, int numKeys, long long* table)   -> table (count, sum,
 *                  errors; the cursor part is left as it is); scratch in/out
 *   GroupbyOffsets(long long* table, int numKeys, void* scratch, long long* offsets,
 *                  long long* groups)
 *   GroupbyReshuffle(int* keys, int* vals, int n, int numKeys, long long* offsets,
 *                    void* scratch, int* outKeys, int* outVals)
 *
 * Keys lie in [0, numKeys]; a key outside is counted as an error and otherwise ignored.
 * numKeys < PTK_GB_MAX_KEYS.
 *
 * table (int64): count[K] | sum[K] | cursor[K] | errors[1], K = numKeys + 1 (the cursor part is
 *   not used by this pipeline; the layout is kept for the oracle's single-pass form).
 * groups (int64): header {G, rows, errors}, then G triples {key, count, sum} in ascending key
 *   order.  offsets[k] = rows with a smaller key (exclusive scan of count).
 *
 * scratch: the rows are cut into chunks of PTK_GB_CHUNK; keys fall into buckets of
 *   PTK_GB_BUCKET consecutive keys (NB = ceil(K / PTK_GB_BUCKET) buckets).
 *   btot  int64[2 NB]      rows and non-empty keys per bucket          (GroupbyMerge)
 *   berr  uint32[B]        invalid keys per chunk                       (GroupbyAccumulate)
 *   boff  uint32[B][NB+1]  per chunk: exclusive scan of its entries by bucket; boff[c][NB] is
 *                          the chunk's number of distinct keys          (GroupbyAccumulate)
 *   entry PTK_GB_ENTRY[B][CHUNK]  per chunk, its distinct keys with their count and sum,
 *                          grouped by bucket (GroupbyAccumulate); GroupbyMerge replaces each
 *                          sum by the entry's base: the first rank of the chunk's rows among
 *                          its key's rows
 *   with B = ceil(n / PTK_GB_CHUNK) chunks.
 */
#ifndef PTK_GROUPBY_H
#define PTK_GROUPBY_H

#if defined(__HIPCC__)
#define PTK_GB_FN static inline __host__ __device__
#else
#define PTK_GB_FN static inline
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define PTK_GB_CHUNK 2048          /* rows per chunk */
#define PTK_GB_BUCKET 256          /* keys per bucket */
#define PTK_GB_MAX_BUCKETS 2048    /* bounds the per-chunk bucket histogram in LDS */
#define PTK_GB_MAX_KEYS (PTK_GB_BUCKET * PTK_GB_MAX_BUCKETS)  /* numKeys + 1 <= this */

typedef struct _ptk_gb_entry {
    int key;
    unsigned int count;
    long long sum;  /* GroupbyMerge replaces it by the base: the rows of the key in the chunks
                       merged before this one, i.e. where the chunk's rows start in the group */
} PTK_GB_ENTRY;

typedef struct _ptk_gb_scratch {
    long long *btot;
    unsigned int *berr, *boff;
    PTK_GB_ENTRY *entry;
    long long chunks, buckets;
} PTK_GB_SCRATCH;

PTK_GB_FN long long ptk_gb_chunks(long long n) { return (n + PTK_GB_CHUNK - 1) / PTK_GB_CHUNK; }
PTK_GB_FN long long ptk_gb_buckets(int numKeys) {
    return ((long long)numKeys + 1 + PTK_GB_BUCKET - 1) / PTK_GB_BUCKET;
}
PTK_GB_FN long long ptk_gb_align16(long long b) { return (b + 15) / 16 * 16; }

/* Byte offsets of the scratch arrays and its total size (a multiple of 16). */
PTK_GB_FN long long ptk_gb_scratch_bytes(long long n, int numKeys) {
    const long long B = ptk_gb_chunks(n), NB = ptk_gb_buckets(numKeys);
    long long o = ptk_gb_align16(16 * NB);          /* btot */
    o = ptk_gb_align16(o + 4 * B);                  /* berr */
    o = ptk_gb_align16(o + 4 * B * (NB + 1));       /* boff */
    return o + 16 * B * PTK_GB_CHUNK;               /* entry */
}

PTK_GB_FN PTK_GB_SCRATCH ptk_gb_scratch(void *scratch, long long n, int numKeys) {
    PTK_GB_SCRATCH s;
    const long long B = ptk_gb_chunks(n), NB = ptk_gb_buckets(numKeys);
    char *p = (char *)scratch;
    long long o = 0;
    s.btot = (long long *)(p + o);
    o = ptk_gb_align16(16 * NB);
    s.berr = (unsigned int *)(p + o);
    o = ptk_gb_align16(o + 4 * B);
    s.boff = (unsigned int *)(p + o);
    o = ptk_gb_align16(o + 4 * B * (NB + 1));
    s.entry = (PTK_GB_ENTRY *)(p + o);
    s.chunks = B;
    s.buckets = NB;
    return s;
}

#ifdef __cplusplus
}
#endif

#endif /* PTK_GROUPBY_H */
