/*
 * oracle/groupby_ref.c -- host-task forms of the seven kernels the reference's groupby graph binds
 * (ptaskUnitTest/graphgroupby.cpp:78-387), so the graph can run on the host accelerator.  TEST
 * INFRASTRUCTURE ONLY (the checker; no product code links it).
 *
 * Same blocks and argument order as the device kernels (ptask_amd/csrc/kernels/groupby_ref.hip),
 * with host memory: hashTable[0] points at a malloc'd record of the sorted unique keys,
 * keyCount[0] is their number and countArray[0] a calloc'd array of per-group counts.  Groups are
 * in ascending key order, so `out` is the sorted input.  Pinned by tests/test_groupby_ref_graph.py:
 * the host graph's output equals numpy's sort of the keys and passes the reference's own check
 * (graphgroupby.cpp:394-419); the GPU graph must equal it bit for bit.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int unique;
    int *keys; /* ascending */
} GbRefTable;

static int cmp_int(const void *a, const void *b) {
    const int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

static int rank_of(const GbRefTable *t, int key) {
    int lo = 0, hi = t->unique - 1;
    while (lo <= hi) {
        const int mid = lo + (hi - lo) / 2;
        if (t->keys[mid] < key) lo = mid + 1;
        else if (t->keys[mid] > key) hi = mid - 1;
        else return mid;
    }
    return -1;
}

/* countUniqueKeysIntInt(keyCount, in, hashTable, n) */
void countUniqueKeysIntInt(unsigned int nargs, void **a) {
    (void)nargs;
    int *keyCount = (int *)a[0];
    const int *in = (const int *)a[1];
    void **hashTable = (void **)a[2];
    const int n = (int)(intptr_t)a[3];
    GbRefTable *t = (GbRefTable *)calloc(1, sizeof(GbRefTable));
    t->keys = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    if (n > 0) memcpy(t->keys, in, sizeof(int) * (size_t)n);
    qsort(t->keys, (size_t)(n > 0 ? n : 0), sizeof(int), cmp_int);
    int u = 0;
    for (int i = 0; i < n; ++i)
        if (u == 0 || t->keys[u - 1] != t->keys[i]) t->keys[u++] = t->keys[i];
    t->unique = u;
    hashTable[0] = t;
    keyCount[0] = u;
}

/* allocateCountArray(countArray, keyCount) */
void allocateCountArray(unsigned int nargs, void **a) {
    (void)nargs;
    int **countArray = (int **)a[0];
    const int *keyCount = (const int *)a[1];
    countArray[0] = (int *)calloc((size_t)(keyCount[0] > 0 ? keyCount[0] : 1), sizeof(int));
}

/* keyCountIntInt(in, n, hashTable, globalCount) */
void keyCountIntInt(unsigned int nargs, void **a) {
    (void)nargs;
    const int *in = (const int *)a[0];
    const int n = (int)(intptr_t)a[1];
    const GbRefTable *t = ((GbRefTable **)a[2])[0];
    int *counts = ((int **)a[3])[0];
    for (int i = 0; i < n; ++i) {
        const int r = rank_of(t, in[i]);
        if (r >= 0) counts[r]++;
    }
}

/* singleBlockPrefixSumIntAdd(globalCount, keyCount): exclusive, in place */
void singleBlockPrefixSumIntAdd(unsigned int nargs, void **a) {
    (void)nargs;
    int *counts = ((int **)a[0])[0];
    const int u = ((const int *)a[1])[0];
    int run = 0;
    for (int i = 0; i < u; ++i) {
        const int c = counts[i];
        counts[i] = run;
        run += c;
    }
}

/* reshuffleIntInt(in, n, out, hashTable, globalCount) */
void reshuffleIntInt(unsigned int nargs, void **a) {
    (void)nargs;
    const int *in = (const int *)a[0];
    const int n = (int)(intptr_t)a[1];
    int *out = (int *)a[2];
    const GbRefTable *t = ((GbRefTable **)a[3])[0];
    int *cursor = ((int **)a[4])[0];
    for (int i = 0; i < n; ++i) {
        const int r = rank_of(t, in[i]);
        if (r >= 0) out[cursor[r]++] = in[i];
    }
}

/* destroyHashTableInt(hashTable) */
void destroyHashTableInt(unsigned int nargs, void **a) {
    (void)nargs;
    void **hashTable = (void **)a[0];
    GbRefTable *t = (GbRefTable *)hashTable[0];
    if (t) {
        free(t->keys);
        free(t);
    }
    hashTable[0] = NULL;
}

/* deallocateCountArray(globalCount) */
void deallocateCountArray(unsigned int nargs, void **a) {
    (void)nargs;
    int **globalCount = (int **)a[0];
    free(globalCount[0]);
    globalCount[0] = NULL;
}
