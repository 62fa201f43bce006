/*
 * ptk_groupby.h -- layouts shared by the groupby task kernels, their host-task forms and the
 * applications that build the groupby graph.
 *
 * The reference graph (ptaskUnitTest/graphgroupby.cpp:106-387) counts keys in a device hash
 * table, prefix-sums the counts and reshuffles the input so that equal keys are contiguous; its
 * kernels are not in the tree.  This pipeline keeps that shape (count, prefix, reshuffle), adds
 * an int32 value column aggregated to an int64 count and sum per key, and orders the rows by a
 * stable two-level radix partition without global atomics:
 *
 *   GroupbyHistogram(int* keys, int n, int numKeys, void* scratch)              -> scratch
 *   GroupbyBucketScan(void* scratch, int n, int numKeys)                         scratch in/out
 *   GroupbyPartition(int* keys, int* vals, int n, int numKeys, void* scratch)    scratch in/out
 *   GroupbyKeyCount(void* scratch, int n, int numKeys)                           scratch in/out
 *   GroupbyKeyScan(void* scratch, int n, int numKeys, long long* offsets)        scratch in/out
 *   GroupbyReshuffle(void* scratch, int n, int numKeys, long long* offsets, int* outKeys,
 *                    int* outVals)                                               scratch in/out
 *   GroupbyTotals(void* scratch, int n, int numKeys, long long* table, long long* groups)
 *
 * Keys lie in [0, numKeys]; a key outside is counted as an error and otherwise ignored.
 * numKeys + 1 <= PTK_GB_MAX_KEYS.
 *
 * Every output is bit-reproducible, including the order of rows inside a group: a group holds its
 * rows in input order (a stable grouping), as the oracle's counting sort does.
 *
 * table (int64): count[K] | sum[K] | cursor[K] | errors[1], K = numKeys + 1 (the cursor part is
 *   not used by this pipeline; the layout is kept for the oracle's single-pass form).
 * groups (int64): header {G, rows, errors}, then G triples {key, count, sum} in ascending key
 *   order.  offsets[k] = rows with a smaller key (exclusive scan of count).
 *
 * scratch: the rows are cut into chunks of PTK_GB_CHUNK; keys fall into buckets of PTK_GB_BUCKET
 *   consecutive keys (NB = ceil(K / PTK_GB_BUCKET) buckets, KP = NB * PTK_GB_BUCKET padded keys).
 *   The partitioned rows (level 1: grouped by bucket, rows of a bucket in input order) are cut,
 *   bucket by bucket, into tiles of PTK_GB_TILE rows; bucket b holds tiles [tstart[b],
 *   tstart[b + 1]), at most TM = ceil(n / PTK_GB_TILE) + NB tiles in all.
 *   head   uint64[4]       [0] invalid keys                             (GroupbyPartition)
 *   btot   uint32[NB]      rows per bucket                              (GroupbyBucketScan)
 *   bstart uint32[NB + 1]  first partitioned row of each bucket, then the valid rows
 *   tstart uint32[NB + 1]  first tile of each bucket, then the tiles   (GroupbyPartition)
 *   tinfo  uint32[TM][4]   per tile: its bucket, first and end row   (GroupbyPartition)
 *   nz     uint32[2 NB]    non-empty keys per half bucket               (GroupbyKeyScan)
 *   berr   uint32[C]       invalid keys per chunk                       (GroupbyHistogram)
 *   hist   uint32[C][NB]   per chunk and bucket: its rows (GroupbyHistogram), then the rows of the
 *                          bucket in the earlier chunks                 (GroupbyBucketScan)
 *   kcnt   uint32[KP]      rows per key                                 (GroupbyKeyScan)
 *   c2     uint32[TM][PTK_GB_BUCKET]  per tile and key (low byte): its rows (GroupbyKeyCount),
 *                          then the key's rows in the bucket's earlier tiles (GroupbyKeyScan)
 *   s2     uint64[TM][PTK_GB_BUCKET]  per tile and key: the sum of its values (GroupbyReshuffle)
 *   tval   int32[n]        partitioned values                           (GroupbyPartition)
 *   tlow   uint8[n + 16]   partitioned keys' low bytes (the bucket is the position's)
 *   with C = ceil(n / PTK_GB_CHUNK) chunks.  A row of key k = b * PTK_GB_BUCKET + l is
 *   partitioned to bstart[b] + hist[c][b] + its rank among chunk c's rows of bucket b (counted a
 *   quarter chunk of PTK_GB_SUB rows at a time), and then, from tile t, placed at offsets[k] +
 *   c2[t][l] + its rank among the tile's rows of key k.
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

#define PTK_GB_CHUNK 16384         /* rows per chunk (histogram, bucket scan) */
#define PTK_GB_SUB 4096            /* rows a GroupbyPartition block ranks at once (a quarter chunk) */
#define PTK_GB_TILE 4096           /* partitioned rows per tile (key count, reshuffle) */
#define PTK_GB_BUCKET 256          /* keys per bucket */
#define PTK_GB_HIST_THREADS 512    /* threads of a GroupbyHistogram block */
#define PTK_GB_COUNT_THREADS 256   /* threads of a GroupbyKeyCount block */
#define PTK_GB_PART_THREADS 512    /* threads of a GroupbyPartition block */
#define PTK_GB_TILE_THREADS 512    /* threads of a GroupbyReshuffle block */
#define PTK_GB_SCAN_THREADS 1024   /* threads of a GroupbyBucketScan block */
#define PTK_GB_KEYSCAN_THREADS 1024 /* threads of a GroupbyKeyScan block (half a bucket's keys) */
#define PTK_GB_TOTALS_THREADS 256 /* threads of a GroupbyTotals block (a quarter bucket's keys) */
#define PTK_GB_MAX_BUCKETS 2048    /* bounds the per-chunk bucket counts in LDS */
#define PTK_GB_MAX_KEYS (PTK_GB_BUCKET * PTK_GB_MAX_BUCKETS)  /* numKeys + 1 <= this */

typedef struct _ptk_gb_scratch {
    unsigned long long *head;
    unsigned int *btot, *bstart, *tstart, *tinfo, *nz, *berr, *hist, *kcnt, *c2;
    unsigned long long *s2;
    int *tval;
    unsigned char *tlow;
    long long chunks, buckets, tilesMax;
} PTK_GB_SCRATCH;

PTK_GB_FN long long ptk_gb_chunks(long long n) { return (n + PTK_GB_CHUNK - 1) / PTK_GB_CHUNK; }
PTK_GB_FN long long ptk_gb_buckets(int numKeys) {
    return ((long long)numKeys + 1 + PTK_GB_BUCKET - 1) / PTK_GB_BUCKET;
}
PTK_GB_FN long long ptk_gb_tiles_max(long long n, int numKeys) {
    return (n + PTK_GB_TILE - 1) / PTK_GB_TILE + ptk_gb_buckets(numKeys);
}
PTK_GB_FN long long ptk_gb_align256(long long b) { return (b + 255) / 256 * 256; }

/* Byte offsets of the scratch arrays, in layout order, and the total size (a multiple of 256). */
PTK_GB_FN void ptk_gb_scratch_offsets(long long n, int numKeys, long long off[14]) {
    const long long C = ptk_gb_chunks(n), NB = ptk_gb_buckets(numKeys), TM = ptk_gb_tiles_max(n, numKeys);
    const long long KP = NB * PTK_GB_BUCKET;
    off[0] = 0;                                              /* head */
    off[1] = 256;                                            /* btot */
    off[2] = ptk_gb_align256(off[1] + 4 * NB);               /* bstart */
    off[3] = ptk_gb_align256(off[2] + 4 * (NB + 1));         /* tstart */
    off[4] = ptk_gb_align256(off[3] + 4 * (NB + 1));         /* tinfo */
    off[5] = ptk_gb_align256(off[4] + 16 * TM);              /* nz */
    off[6] = ptk_gb_align256(off[5] + 8 * NB);               /* berr */
    off[7] = ptk_gb_align256(off[6] + 4 * C);                /* hist */
    off[8] = ptk_gb_align256(off[7] + 4 * C * NB);           /* kcnt */
    off[9] = ptk_gb_align256(off[8] + 4 * KP);               /* c2 */
    off[10] = ptk_gb_align256(off[9] + 4 * TM * PTK_GB_BUCKET);   /* s2 */
    off[11] = ptk_gb_align256(off[10] + 8 * TM * PTK_GB_BUCKET);  /* tval */
    off[12] = ptk_gb_align256(off[11] + 4 * n);              /* tlow */
    off[13] = ptk_gb_align256(off[12] + n + 16);             /* total */
}

PTK_GB_FN long long ptk_gb_scratch_bytes(long long n, int numKeys) {
    long long off[14];
    ptk_gb_scratch_offsets(n, numKeys, off);
    return off[13];
}

PTK_GB_FN PTK_GB_SCRATCH ptk_gb_scratch(void *scratch, long long n, int numKeys) {
    PTK_GB_SCRATCH s;
    long long off[14];
    ptk_gb_scratch_offsets(n, numKeys, off);
    char *p = (char *)scratch;
    s.head = (unsigned long long *)(p + off[0]);
    s.btot = (unsigned int *)(p + off[1]);
    s.bstart = (unsigned int *)(p + off[2]);
    s.tstart = (unsigned int *)(p + off[3]);
    s.tinfo = (unsigned int *)(p + off[4]);
    s.nz = (unsigned int *)(p + off[5]);
    s.berr = (unsigned int *)(p + off[6]);
    s.hist = (unsigned int *)(p + off[7]);
    s.kcnt = (unsigned int *)(p + off[8]);
    s.c2 = (unsigned int *)(p + off[9]);
    s.s2 = (unsigned long long *)(p + off[10]);
    s.tval = (int *)(p + off[11]);
    s.tlow = (unsigned char *)(p + off[12]);
    s.chunks = ptk_gb_chunks(n);
    s.buckets = ptk_gb_buckets(numKeys);
    s.tilesMax = ptk_gb_tiles_max(n, numKeys);
    return s;
}

#ifdef __cplusplus
}
#endif

#endif /* PTK_GROUPBY_H */
