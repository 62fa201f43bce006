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
 *   GroupbyMerge(void* scratch, int n, int numKeys)   scratch in/out.  Blocks of exactly
 *                  PTK_GB_BUCKET threads.
 *   GroupbyTotals(void* scratch, int n, int numKeys, long long* table)   -> table (count, sum,
 *                  errors; the cursor part is left as it is); scratch in/out.  Blocks of
 *                  exactly PTK_GB_BUCKET threads.
 *   GroupbyOffsets(long long* table, int numKeys, void* scratch, long long* offsets,
 *                  long long* groups)
 *   GroupbyReshuffle(int* keys, int* vals, int n, int numKeys, long long* offsets,
 *                    void* scratch, int* outKeys, int* outVals)   Blocks of exactly
 *                    PTK_GB_PLACE_THREADS threads.
 *
 * Keys lie in [0, numKeys]; a key outside is counted as an error and otherwise ignored.
 * numKeys < PTK_GB_MAX_KEYS.
 *
 * Every output is bit-reproducible, including the order of rows inside a group: a group holds its
 * rows in input order (a stable grouping), as the oracle's counting sort does.
 *
 * table (int64): count[K] | sum[K] | cursor[K] | errors[1], K = numKeys + 1 (the cursor part is
 *   not used by this pipeline; the layout is kept for the oracle's single-pass form).
 * groups (int64): header {G, rows, errors}, then G triples {key, count, sum} in ascending key
 *   order.  offsets[k] = rows with a smaller key (exclusive scan of count).
 *
 * scratch: the rows are cut into chunks of PTK_GB_CHUNK, the chunks into merge groups of
 *   PTK_GB_GROUP consecutive chunks; keys fall into buckets of PTK_GB_BUCKET consecutive keys
 *   (NB = ceil(K / PTK_GB_BUCKET) buckets, KP = NB * PTK_GB_BUCKET padded keys).
 *   btot  int64[2 NB]      rows and non-empty keys per bucket          (GroupbyTotals)
 *   berr  uint32[B]        invalid keys per chunk                       (GroupbyAccumulate)
 *   boff  uint32[B][NB+1]  per chunk: exclusive scan of its entries by bucket; boff[c][NB] is
 *                          the chunk's number of distinct keys          (GroupbyAccumulate)
 *   gsum  uint64[NG][KP]   per merge group and key: the sum of its values in the group
 *   gcnt  uint32[NG][KP]   per merge group and key: its rows in the group (GroupbyMerge), then
 *                          its rows in the earlier groups (GroupbyTotals)
 *   entry PTK_GB_ENTRY[B][CHUNK]  per chunk, its distinct keys with their count and sum,
 *                          grouped by bucket (GroupbyAccumulate); GroupbyMerge replaces each
 *                          sum by the entry's base inside its merge group: the key's rows in the
 *                          group's earlier chunks
 *   with B = ceil(n / PTK_GB_CHUNK) chunks and NG = ceil(B / PTK_GB_GROUP) merge groups.  A row
 *   of key k in chunk c is therefore placed at offsets[k] + gcnt[c / PTK_GB_GROUP][k] + the
 *   entry's base + its rank among the chunk's rows of key k.
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
#define PTK_GB_GROUP 256           /* chunks per merge group */
#define PTK_GB_PLACE_THREADS 512   /* threads of a GroupbyReshuffle block (8 waves) */
#define PTK_GB_MAX_BUCKETS 2048    /* bounds the per-chunk bucket histogram in LDS */
#define PTK_GB_MAX_KEYS (PTK_GB_BUCKET * PTK_GB_MAX_BUCKETS)  /* numKeys + 1 <= this */

typedef struct _ptk_gb_entry {
    int key;
    unsigned int count;
    long long sum;  /* GroupbyMerge replaces it by the base: the rows of the key in the chunks
                       of the same merge group before this one */
} PTK_GB_ENTRY;

typedef struct _ptk_gb_scratch {
    long long *btot;
    unsigned int *berr, *boff, *gcnt;
    unsigned long long *gsum;
    PTK_GB_ENTRY *entry;
    long long chunks, buckets, groups;
} PTK_GB_SCRATCH;

PTK_GB_FN long long ptk_gb_chunks(long long n) { return (n + PTK_GB_CHUNK - 1) / PTK_GB_CHUNK; }
PTK_GB_FN long long ptk_gb_buckets(int numKeys) {
    return ((long long)numKeys + 1 + PTK_GB_BUCKET - 1) / PTK_GB_BUCKET;
}
PTK_GB_FN long long ptk_gb_groups(long long n) { return (ptk_gb_chunks(n) + PTK_GB_GROUP - 1) / PTK_GB_GROUP; }
PTK_GB_FN long long ptk_gb_align16(long long b) { return (b + 15) / 16 * 16; }

/* Byte offsets of the scratch arrays, in layout order, and the total size (a multiple of 16).
 * btot comes first, so its offset does not depend on n. */
PTK_GB_FN void ptk_gb_scratch_offsets(long long n, int numKeys, long long off[7]) {
    const long long B = ptk_gb_chunks(n), NB = ptk_gb_buckets(numKeys), NG = ptk_gb_groups(n);
    const long long KP = NB * PTK_GB_BUCKET;
    off[0] = 0;                                     /* btot */
    off[1] = ptk_gb_align16(16 * NB);               /* berr */
    off[2] = ptk_gb_align16(off[1] + 4 * B);        /* boff */
    off[3] = ptk_gb_align16(off[2] + 4 * B * (NB + 1));  /* gsum */
    off[4] = ptk_gb_align16(off[3] + 8 * NG * KP);  /* gcnt */
    off[5] = ptk_gb_align16(off[4] + 4 * NG * KP);  /* entry */
    off[6] = off[5] + 16 * B * PTK_GB_CHUNK;        /* total */
}

PTK_GB_FN long long ptk_gb_scratch_bytes(long long n, int numKeys) {
    long long off[7];
    ptk_gb_scratch_offsets(n, numKeys, off);
    return off[6];
}

PTK_GB_FN PTK_GB_SCRATCH ptk_gb_scratch(void *scratch, long long n, int numKeys) {
    PTK_GB_SCRATCH s;
    long long off[7];
    ptk_gb_scratch_offsets(n, numKeys, off);
    char *p = (char *)scratch;
    s.btot = (long long *)(p + off[0]);
    s.berr = (unsigned int *)(p + off[1]);
    s.boff = (unsigned int *)(p + off[2]);
    s.gsum = (unsigned long long *)(p + off[3]);
    s.gcnt = (unsigned int *)(p + off[4]);
    s.entry = (PTK_GB_ENTRY *)(p + off[5]);
    s.chunks = ptk_gb_chunks(n);
    s.buckets = ptk_gb_buckets(numKeys);
    s.groups = ptk_gb_groups(n);
    return s;
}

#ifdef __cplusplus
}
#endif

#endif /* PTK_GROUPBY_H */
