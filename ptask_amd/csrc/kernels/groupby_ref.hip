// groupby_ref.hip -- the seven task kernels the reference's groupby graph binds by name
// (ptaskUnitTest/graphgroupby.cpp:78-387): countUniqueKeysIntInt, allocateCountArray,
// keyCountIntInt, singleBlockPrefixSumIntAdd, reshuffleIntInt, destroyHashTableInt and
// deallocateCountArray, with the argument order of the ports' formal indices there.  The reference
// kernels' source is not in its tree; what they must compute is fixed by the graph and its check
// (graphgroupby.cpp:394-419): `out` holds the input keys with equal keys contiguous.  Here the
// groups come in ascending key order, so `out` is the sorted input, the same on every run.
//
// The reference builds a hash table of device-malloc'd nodes under 256 bucket pointers.  Here the
// state lives in a bump heap inside this code object (g_gbrHeap, allocated once per process when
// the code object is loaded) and the application's blocks carry only addresses and sizes:
//   hashTable  (256 x 8 bytes, NULL on entry)  [0] table address (or a busy / failed mark),
//              [1] capacity, [2] epoch; the rest stays unused
//   keyCount   (10 ints, zero on entry)        [0] unique keys, [1] error (heap exhausted),
//              [2..3] table address, [4] capacity, [5] epoch, [6] least key, [7] greatest key
//   countArray (one pointer)                   address of the per-group counts / cursors
// The table is open addressing with linear probing over cap >= 2n slots of 64-bit tags
// (epoch << 32 | key); a slot whose epoch is not this table's is empty, so a table never needs
// clearing.  A second cap-long int array holds each occupied slot's group rank (the number of
// smaller keys), which allocateCountArray computes from a bitmap over [least, greatest] key (range
// up to 2^24) or by a radix sort of the unique keys.  The heap is first fit over a sorted list of
// live allocations under a lock one thread takes, so runs in flight at once (a run's clean-up
// tasks may run after the next run's first task) each hold their own memory.
// Capacity: GBR_HEAP_BYTES bounds the runs' tables (12 bytes per slot, cap >= 2n) plus counts;
// one run of about 8M rows fits.
#include <hip/hip_runtime.h>
#include <climits>
#include <cstdint>

#define GBR_HEAP_BYTES (320ull << 20)

namespace {

constexpr unsigned long long kBusy = 1, kFailed = 3;  // table-word marks; real addresses are 256-aligned
constexpr unsigned kMaxAllocs = 64;  // live heap allocations (two per run in flight)
constexpr unsigned long long kNotFound = ~0ull;
enum : int { HT_TABLE = 0, HT_CAP = 1, HT_EPOCH = 2 };
enum : int { KC_UNIQUE = 0, KC_ERROR = 1, KC_TABLE_LO = 2, KC_TABLE_HI = 3, KC_CAP = 4, KC_EPOCH = 5, KC_MIN = 6, KC_MAX = 7 };
constexpr unsigned long long kCountHeader = 64;  // bytes before the counts in their allocation
constexpr int kMaxThreads = 1024;

}  // namespace

__device__ unsigned int g_gbrLock;             // 1 while one thread edits the allocation list
__device__ unsigned int g_gbrAllocCount;       // live allocations, sorted by offset below
__device__ unsigned long long g_gbrAllocOff[kMaxAllocs], g_gbrAllocLen[kMaxAllocs];
__device__ unsigned int g_gbrEpoch;
__device__ __attribute__((aligned(256))) unsigned char g_gbrHeap[GBR_HEAP_BYTES];

namespace {

__device__ inline unsigned long long LoadAgent(const unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned LoadAgent(const unsigned *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void StoreAgent(unsigned long long *p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The allocator's lock: taken by a single thread of a kernel (never by several lanes of one wave),
// held for a walk of at most kMaxAllocs entries.
__device__ void HeapLock() {
    unsigned expected = 0;
    while (!__hip_atomic_compare_exchange_strong(&g_gbrLock, &expected, 1u, __ATOMIC_ACQUIRE, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)) {
        expected = 0;
        __builtin_amdgcn_s_sleep(2);
    }
}
__device__ void HeapUnlock() { __hip_atomic_store(&g_gbrLock, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT); }

// First fit over the live allocations (sorted by offset); nullptr when no gap is large enough or
// kMaxAllocs allocations are live.
__device__ unsigned char *HeapAlloc(unsigned long long bytes) {
    bytes = (bytes + 255) & ~255ull;
    HeapLock();
    const unsigned n = LoadAgent(&g_gbrAllocCount);
    unsigned long long end = 0;
    unsigned pos = n;
    for (unsigned i = 0; i < n; ++i) {
        const unsigned long long off = LoadAgent(&g_gbrAllocOff[i]);
        if (off - end >= bytes) {
            pos = i;
            break;
        }
        end = off + LoadAgent(&g_gbrAllocLen[i]);
    }
    unsigned char *p = nullptr;
    if (n < kMaxAllocs && (pos < n || GBR_HEAP_BYTES - end >= bytes)) {
        for (unsigned i = n; i > pos; --i) {
            StoreAgent(&g_gbrAllocOff[i], LoadAgent(&g_gbrAllocOff[i - 1]));
            StoreAgent(&g_gbrAllocLen[i], LoadAgent(&g_gbrAllocLen[i - 1]));
        }
        StoreAgent(&g_gbrAllocOff[pos], end);
        StoreAgent(&g_gbrAllocLen[pos], bytes);
        __hip_atomic_store(&g_gbrAllocCount, n + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        p = g_gbrHeap + end;
    }
    HeapUnlock();
    return p;
}

// Ends the allocation that starts at p (an unknown address is ignored).
__device__ void HeapFree(const void *p) {
    const unsigned long long off = (unsigned long long)((const unsigned char *)p - g_gbrHeap);
    HeapLock();
    const unsigned n = LoadAgent(&g_gbrAllocCount);
    for (unsigned i = 0; i < n; ++i) {
        if (LoadAgent(&g_gbrAllocOff[i]) != off) continue;
        for (unsigned j = i; j + 1 < n; ++j) {
            StoreAgent(&g_gbrAllocOff[j], LoadAgent(&g_gbrAllocOff[j + 1]));
            StoreAgent(&g_gbrAllocLen[j], LoadAgent(&g_gbrAllocLen[j + 1]));
        }
        __hip_atomic_store(&g_gbrAllocCount, n - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
    }
    HeapUnlock();
}

// lowbias32 (C. Wellons' integer hash): spreads clustered keys over the table
__device__ inline unsigned Hash(int key) {
    unsigned x = (unsigned)key;
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ inline unsigned long long Tag(unsigned ep, int key) { return ((unsigned long long)ep << 32) | (unsigned)key; }

// Slot of a key in a complete table, kNotFound if absent (an empty slot ends the probe; at least
// half the slots are empty).
__device__ inline unsigned long long Find(const unsigned long long *slots, unsigned long long cap, unsigned ep, int key) {
    const unsigned long long tag = Tag(ep, key);
    unsigned long long h = Hash(key) & (cap - 1);
    for (;;) {
        const unsigned long long v = slots[h];
        if (v == tag) return h;
        if ((unsigned)(v >> 32) != ep) return kNotFound;
        h = (h + 1) & (cap - 1);
    }
}

__device__ inline int FlatThread() { return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z); }
__device__ inline int FlatThreads() { return blockDim.x * blockDim.y * blockDim.z; }
__device__ inline long long FlatBlock() {
    return blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
}
__device__ inline long long FlatBlocks() { return (long long)gridDim.x * gridDim.y * gridDim.z; }

__device__ inline bool LiveTable(unsigned long long v) { return v != 0 && v != kBusy && v != kFailed; }

// Exclusive scan of one value per thread over the block (at most kMaxThreads threads); returns the
// thread's exclusive prefix and sets *total.  Every thread of the block must call it.
__device__ int BlockScan(int v, int *buf, int *total) {
    const int t = FlatThread(), nt = FlatThreads();
    buf[t] = v;
    __syncthreads();
    for (int o = 1; o < nt; o <<= 1) {
        const int add = t >= o ? buf[t - o] : 0;
        __syncthreads();
        buf[t] += add;
        __syncthreads();
    }
    const int incl = buf[t];
    *total = buf[nt - 1];
    __syncthreads();
    return incl - v;
}

// Rows of a wave that share a group take one atomic between them: the lowest lane of each group
// adds the group's row count to counts[r] and hands the old value to its peers.  Returns the
// position of this lane's row among the group's rows so far (base + rank among peers), or -1 for
// a lane with r < 0.  Every lane of the wave must call it.
__device__ inline long long WaveAggregatedAdd(int *counts, int r) {
    const int lane = __lane_id();
    unsigned long long active = __ballot(r >= 0);
    long long mine = -1;
    while (active) {
        const int leader = __ffsll((long long)active) - 1;
        const int lr = __shfl(r, leader, 64);
        const unsigned long long peers = __ballot(r == lr) & active;
        int base = 0;
        if (lane == leader) base = atomicAdd(&counts[lr], __popcll(peers));
        base = __shfl(base, leader, 64);
        if ((peers >> lane) & 1ull) mine = base + __popcll(peers & ((1ull << lane) - 1ull));
        active &= ~peers;
    }
    return mine;
}

}  // namespace

// countUniqueKeysIntInt(keyCount, in, hashTable, n)  (graphgroupby.cpp:332-387)
// The first block to arrive allocates the table; the others wait for its address.  Each row's key
// is inserted unless present; keyCount[0] counts the insertions, [6] and [7] the key range.
extern "C" __global__ void __launch_bounds__(kMaxThreads)
countUniqueKeysIntInt(int *keyCount, const int *in, unsigned long long *hashTable, int n) {
    __shared__ unsigned long long sTable, sCap;
    __shared__ unsigned sEpoch;
    __shared__ int sFresh, sMin, sMax;
    const int t = FlatThread(), nt = FlatThreads();
    if (t == 0) {
        unsigned long long v = __hip_atomic_load(&hashTable[HT_TABLE], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (v == 0) {
            v = atomicCAS(&hashTable[HT_TABLE], 0ull, kBusy);
            if (v == 0) {
                unsigned long long cap = 64;
                while (cap < 2ull * (unsigned long long)(n > 0 ? n : 0)) cap <<= 1;
                unsigned char *p = cap <= GBR_HEAP_BYTES / 12 ? HeapAlloc(cap * 12) : nullptr;
                const unsigned ep = atomicAdd(&g_gbrEpoch, 1u) + 1u;
                hashTable[HT_CAP] = cap;
                hashTable[HT_EPOCH] = ep;
                keyCount[KC_ERROR] = p ? 0 : 1;
                keyCount[KC_TABLE_LO] = (int)(unsigned)(uintptr_t)p;
                keyCount[KC_TABLE_HI] = (int)(unsigned)((uintptr_t)p >> 32);
                keyCount[KC_CAP] = (int)(p ? cap : 0);
                keyCount[KC_EPOCH] = (int)ep;
                keyCount[KC_MIN] = INT_MAX;
                keyCount[KC_MAX] = INT_MIN;
                v = p ? (unsigned long long)(uintptr_t)p : kFailed;
                __hip_atomic_store(&hashTable[HT_TABLE], v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        while (v == kBusy) {
            __builtin_amdgcn_s_sleep(4);
            v = __hip_atomic_load(&hashTable[HT_TABLE], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        }
        // the word no longer changes during this kernel; an acquiring load orders the reads of
        // the capacity and epoch after the allocating block's release
        v = __hip_atomic_load(&hashTable[HT_TABLE], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        sTable = v;
        sCap = LoadAgent(&hashTable[HT_CAP]);
        sEpoch = (unsigned)LoadAgent(&hashTable[HT_EPOCH]);
        sFresh = 0;
        sMin = INT_MAX;
        sMax = INT_MIN;
    }
    __syncthreads();
    if (!LiveTable(sTable)) return;
    unsigned long long *slots = (unsigned long long *)sTable;
    const unsigned long long cap = sCap;
    const unsigned ep = sEpoch;
    int fresh = 0, kmin = INT_MAX, kmax = INT_MIN;
    for (long long i = FlatBlock() * nt + t; i < n; i += FlatBlocks() * nt) {
        const int key = in[i];
        kmin = min(kmin, key);
        kmax = max(kmax, key);
        const unsigned long long tag = Tag(ep, key);
        unsigned long long h = Hash(key) & (cap - 1);
        for (;;) {
            const unsigned long long cur = LoadAgent(&slots[h]);
            if (cur == tag) break;
            if ((unsigned)(cur >> 32) == ep) {  // another key's slot
                h = (h + 1) & (cap - 1);
                continue;
            }
            if (atomicCAS(&slots[h], cur, tag) == cur) {
                ++fresh;
                break;
            }
            // another row claimed this empty slot first: look at it again
        }
    }
    if (fresh) atomicAdd(&sFresh, fresh);
    atomicMin(&sMin, kmin);
    atomicMax(&sMax, kmax);
    __syncthreads();
    if (t == 0) {
        if (sFresh) atomicAdd(&keyCount[KC_UNIQUE], sFresh);
        if (sMin != INT_MAX) atomicMin(&keyCount[KC_MIN], sMin);
        if (sMax != INT_MIN) atomicMax(&keyCount[KC_MAX], sMax);
    }
}

// allocateCountArray(countArray, keyCount)  (graphgroupby.cpp:297-330)
// Allocates the per-group counts (zeroed) and ranks every occupied slot by key.  One block works
// (the reference launches one); others return.
extern "C" __global__ void __launch_bounds__(kMaxThreads) allocateCountArray(int **countArray, const int *keyCount) {
    if (FlatBlock() != 0) return;
    __shared__ unsigned char *sBase;
    __shared__ int sBuf[kMaxThreads];
    const int t = FlatThread(), nt = FlatThreads();
    const int U = keyCount[KC_UNIQUE];
    const unsigned long long table =
        ((unsigned long long)(unsigned)keyCount[KC_TABLE_HI] << 32) | (unsigned)keyCount[KC_TABLE_LO];
    const unsigned long long cap = (unsigned)keyCount[KC_CAP];
    const unsigned ep = (unsigned)keyCount[KC_EPOCH];
    const int kmin = keyCount[KC_MIN], kmax = keyCount[KC_MAX];
    const bool usable = keyCount[KC_ERROR] == 0 && table != 0 && cap != 0 && U >= 0;
    const unsigned long long range = U > 0 ? (unsigned long long)((long long)kmax - (long long)kmin) + 1 : 0;
    const bool dense = range <= (1ull << 24);
    const unsigned long long words = dense ? (range + 31) / 32 : 0;
    if (t == 0) {
        const unsigned long long bytes = kCountHeader + 4ull * U + (dense ? 8ull * words : 8ull * U);
        sBase = usable ? HeapAlloc(bytes) : nullptr;
        if (sBase) ((int *)sBase)[0] = U;
        countArray[0] = sBase ? (int *)(sBase + kCountHeader) : nullptr;
    }
    __syncthreads();
    if (!sBase) return;
    int *counts = (int *)(sBase + kCountHeader);
    const unsigned long long *slots = (const unsigned long long *)table;
    int *rank = (int *)(slots + cap);
    for (int i = t; i < U; i += nt) counts[i] = 0;
    if (dense) {
        unsigned *bits = (unsigned *)(counts + U), *pre = bits + words;
        for (unsigned long long w = t; w < words; w += nt) bits[w] = 0;
        __syncthreads();
        for (unsigned long long s = t; s < cap; s += nt) {
            const unsigned long long v = slots[s];
            if ((unsigned)(v >> 32) != ep) continue;
            const unsigned off = (unsigned)((long long)(int)(unsigned)v - kmin);
            atomicOr(&bits[off >> 5], 1u << (off & 31));
        }
        __syncthreads();
        int carry = 0;
        for (unsigned long long base = 0; base < words; base += nt) {
            const unsigned long long w = base + t;
            const int c = w < words ? __popc(LoadAgent(&bits[w])) : 0;
            int total;
            const int ex = BlockScan(c, sBuf, &total);
            if (w < words) pre[w] = (unsigned)(carry + ex);
            carry += total;
        }
        __syncthreads();
        for (unsigned long long s = t; s < cap; s += nt) {
            const unsigned long long v = slots[s];
            if ((unsigned)(v >> 32) != ep) continue;
            const unsigned off = (unsigned)((long long)(int)(unsigned)v - kmin);
            const unsigned below = LoadAgent(&bits[off >> 5]) & ((1u << (off & 31)) - 1u);
            rank[s] = (int)(pre[off >> 5] + __popc(below));
        }
        return;
    }
    // a wide key range: compact the unique keys in slot order, radix-sort them (one bit per pass,
    // stable splits, 32 passes so the result ends where it began), then rank by sorted position
    int *keys = counts + U, *tmp = keys + U;
    int carry = 0;
    for (unsigned long long base = 0; base < cap; base += nt) {
        const unsigned long long s = base + t;
        const unsigned long long v = s < cap ? slots[s] : 0;
        const int f = s < cap && (unsigned)(v >> 32) == ep ? 1 : 0;
        int total;
        const int ex = BlockScan(f, sBuf, &total);
        if (f) keys[carry + ex] = (int)(unsigned)v;
        carry += total;
    }
    __syncthreads();
    for (int bit = 0; bit < 32; ++bit) {
        const int *src = bit & 1 ? tmp : keys;
        int *dst = bit & 1 ? keys : tmp;
        int zeros = 0;
        for (int base = 0; base < U; base += nt) {
            const int i = base + t;
            const int z = i < U && !((((unsigned)src[i] ^ 0x80000000u) >> bit) & 1u) ? 1 : 0;
            int total;
            BlockScan(z, sBuf, &total);
            zeros += total;
        }
        int c0 = 0, c1 = 0;
        for (int base = 0; base < U; base += nt) {
            const int i = base + t;
            const int k = i < U ? src[i] : 0;
            const int z = i < U && !((((unsigned)k ^ 0x80000000u) >> bit) & 1u) ? 1 : 0;
            int tz;
            const int ez = BlockScan(z, sBuf, &tz);
            if (i < U) dst[z ? c0 + ez : zeros + c1 + (t - ez)] = k;
            c0 += tz;
            c1 += min(nt, U - base) - tz;
        }
        __syncthreads();
    }
    for (int r = t; r < U; r += nt) {
        const unsigned long long s = Find(slots, cap, ep, keys[r]);
        if (s != kNotFound) rank[s] = r;
    }
}

// keyCountIntInt(in, n, hashTable, globalCount)  (graphgroupby.cpp:245-295): rows per group
extern "C" __global__ void __launch_bounds__(kMaxThreads)
keyCountIntInt(const int *in, int n, const unsigned long long *hashTable, int **globalCount) {
    int *counts = globalCount[0];
    const unsigned long long table = hashTable[HT_TABLE];
    if (!counts || !LiveTable(table)) return;
    const unsigned long long cap = hashTable[HT_CAP];
    const unsigned ep = (unsigned)hashTable[HT_EPOCH];
    const unsigned long long *slots = (const unsigned long long *)table;
    const int *rank = (const int *)(slots + cap);
    const int t = FlatThread(), nt = FlatThreads();
    for (long long base = FlatBlock() * nt; base < n; base += FlatBlocks() * nt) {
        const long long i = base + t;
        int r = -1;
        if (i < n) {
            const unsigned long long s = Find(slots, cap, ep, in[i]);
            if (s != kNotFound) r = rank[s];
        }
        WaveAggregatedAdd(counts, r);
    }
}

// singleBlockPrefixSumIntAdd(globalCount, keyCount)  (graphgroupby.cpp:204-242): exclusive
// prefix sum of the counts in place, by one block (the reference launches one); others return.
extern "C" __global__ void __launch_bounds__(kMaxThreads) singleBlockPrefixSumIntAdd(int **globalCount, const int *keyCount) {
    if (FlatBlock() != 0) return;
    __shared__ int sBuf[kMaxThreads];
    int *counts = globalCount[0];
    if (!counts) return;
    const int U = keyCount[KC_UNIQUE];
    const int t = FlatThread(), nt = FlatThreads();
    int carry = 0;
    for (int base = 0; base < U; base += nt) {
        const int i = base + t;
        const int c = i < U ? counts[i] : 0;
        int total;
        const int ex = BlockScan(c, sBuf, &total);
        if (i < U) counts[i] = carry + ex;
        carry += total;
    }
}

// reshuffleIntInt(in, n, out, hashTable, globalCount)  (graphgroupby.cpp:147-201): every row's
// key to the next free position of its group
extern "C" __global__ void __launch_bounds__(kMaxThreads)
reshuffleIntInt(const int *in, int n, int *out, unsigned long long *hashTable, int **globalCount) {
    int *cursor = globalCount[0];
    const unsigned long long table = hashTable[HT_TABLE];
    if (!cursor || !LiveTable(table)) return;
    const unsigned long long cap = hashTable[HT_CAP];
    const unsigned ep = (unsigned)hashTable[HT_EPOCH];
    const unsigned long long *slots = (const unsigned long long *)table;
    const int *rank = (const int *)(slots + cap);
    const int t = FlatThread(), nt = FlatThreads();
    for (long long base = FlatBlock() * nt; base < n; base += FlatBlocks() * nt) {
        const long long i = base + t;
        int r = -1, key = 0;
        if (i < n) {
            key = in[i];
            const unsigned long long s = Find(slots, cap, ep, key);
            if (s != kNotFound) r = rank[s];
        }
        const long long pos = WaveAggregatedAdd(cursor, r);
        if (pos >= 0) out[pos] = key;
    }
}

// destroyHashTableInt(hashTable)  (graphgroupby.cpp:106-144): the table's memory goes back
extern "C" __global__ void destroyHashTableInt(unsigned long long *hashTable) {
    if (FlatBlock() != 0 || FlatThread() != 0) return;
    const unsigned long long v = hashTable[HT_TABLE];
    if (LiveTable(v)) HeapFree((const void *)v);
    hashTable[HT_TABLE] = 0;
}

// deallocateCountArray(globalCount)  (graphgroupby.cpp:76-104)
extern "C" __global__ void deallocateCountArray(int **globalCount) {
    if (FlatBlock() != 0 || FlatThread() != 0) return;
    if (globalCount[0]) HeapFree((const unsigned char *)globalCount[0] - kCountHeader);
    globalCount[0] = nullptr;
}
