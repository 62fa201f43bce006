// scan.hip -- exclusive prefix sums as PTask task kernels (code object scan.hsaco).
//
// Semantics (reference oracle host_pfxsum, ptaskUnitTest/pfxsum.cpp:56-69): out[0] = 0,
// out[i] = out[i-1] + in[i-1], in two's-complement int32 (wrapping) arithmetic.
//
//   pfxsum(int* pin, int* pout, int N)           -- the reference task ABI (pfxsum.cu:21-25,
//       formal indices pin 0, pout 1, N 2 as cpfxsum.cpp wires them).  Any N >= 0: one block
//       walks the array in chunks of 2*blockDim with a running carry; extra blocks exit.  The
//       reference kernel is limited to 256 elements (pfxsum.cu:6); this one is not.
//   PfxScanReduce(const int* pin, int N, int* tileSums)          -- large N, pass 1: one total
//       per tile of PFX_TILE elements.
//   PfxScanApply(const int* pin, int N, const int* tileSums, int* pout) -- pass 2: each block
//       sums the totals before its tile, then scans its tile.  12 B of HBM traffic per element.
//   *64 variants: int64 elements.
//
// Wave-level scans use DPP-free shuffles (__shfl_up over 64-lane wavefronts), then one LDS
// exchange of the per-wave totals.  Integer addition is associative, so the result is exact
// regardless of the reduction order.
#include <hip/hip_runtime.h>

#define PFX_TILE 4096  // elements per tile of the large scan (1024 threads x 4)

template <typename T>
__device__ inline T wave_inclusive(T v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
    }
    return v;
}

template <typename T>
__device__ inline T wave_total(T v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Block-wide exclusive scan of one value per thread (blockDim.x a multiple of 64, <= 1024).
// Returns the exclusive prefix; *total receives the block total.
template <typename T>
__device__ inline T block_exclusive(T v, T *total) {
    __shared__ T waveSums[16];
    __shared__ T blockTotal;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    T inc = wave_inclusive(v);
    if (lane == 63 || (int)threadIdx.x == (int)blockDim.x - 1) waveSums[wave] = inc;
    __syncthreads();
    if (wave == 0) {
        T s = lane < nw ? waveSums[lane] : T(0);
        T si = wave_inclusive(s);
        if (lane < nw) waveSums[lane] = si - s;  // exclusive prefix of the wave totals
        if (lane == nw - 1) blockTotal = si;
    }
    __syncthreads();
    T res = waveSums[wave] + inc - v;
    *total = blockTotal;
    __syncthreads();  // the shared slots are reused by the next call
    return res;
}

// One block scans the whole array, two elements per thread per chunk, carrying the running sum.
// Launch contract of every kernel here: blockDim.x a multiple of 64 (whole wavefronts), at most
// 1024, blockDim.y == blockDim.z == 1.
__device__ inline bool block_ok() { return (blockDim.x & 63) == 0 && blockDim.y == 1 && blockDim.z == 1; }

template <typename T>
__device__ inline void single_block_scan(const T *__restrict__ in, T *__restrict__ out, long long n) {
    if (!block_ok() || blockIdx.x != 0 || blockIdx.y != 0 || blockIdx.z != 0) return;
    T carry = 0;
    const long long chunk = 2ll * blockDim.x;
    for (long long base = 0; base < n; base += chunk) {
        const long long i0 = base + 2ll * threadIdx.x, i1 = i0 + 1;
        const T a = i0 < n ? in[i0] : T(0), b = i1 < n ? in[i1] : T(0);
        T tot;
        T ex = block_exclusive<T>(a + b, &tot);
        if (i0 < n) out[i0] = carry + ex;
        if (i1 < n) out[i1] = carry + ex + a;
        carry += tot;
    }
}

// int32 with wrapping adds: compute in unsigned arithmetic (defined overflow), store the bits.
extern "C" __global__ void __launch_bounds__(1024) pfxsum(int *pin, int *pout, int N) {
    single_block_scan<unsigned>((const unsigned *)pin, (unsigned *)pout, N < 0 ? 0 : N);
}

extern "C" __global__ void __launch_bounds__(1024) pfxsum64(long long *pin, long long *pout, int N) {
    single_block_scan<unsigned long long>((const unsigned long long *)pin, (unsigned long long *)pout,
                                          N < 0 ? 0 : N);
}

template <typename T>
__device__ inline void tile_reduce(const T *__restrict__ in, long long n, T *__restrict__ tileSums) {
    if (!block_ok()) return;
    const long long ntiles = (n + PFX_TILE - 1) / PFX_TILE;
    for (long long t = blockIdx.x; t < ntiles; t += gridDim.x) {
        T s = 0;
        const long long base = t * PFX_TILE;
        for (int e = threadIdx.x; e < PFX_TILE; e += blockDim.x) {
            const long long i = base + e;
            if (i < n) s += in[i];
        }
        s = wave_total(s);
        __shared__ T part[16];
        if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            T tot = 0;
            for (int w = 0; w < (int)((blockDim.x + 63) >> 6); ++w) tot += part[w];
            tileSums[t] = tot;
        }
        __syncthreads();
    }
}

template <typename T>
__device__ inline void tile_apply(const T *__restrict__ in, long long n, const T *__restrict__ tileSums,
                                  T *__restrict__ out) {
    if (!block_ok()) return;
    const long long ntiles = (n + PFX_TILE - 1) / PFX_TILE;
    for (long long t = blockIdx.x; t < ntiles; t += gridDim.x) {
        // offset of this tile: sum of the totals before it
        T pre = 0;
        for (long long u = threadIdx.x; u < t; u += blockDim.x) pre += tileSums[u];
        T dummy;
        T preEx = block_exclusive<T>(pre, &dummy);
        (void)preEx;
        T offset = dummy;
        // scan the tile: each thread owns PFX_TILE/blockDim consecutive elements
        const int per = PFX_TILE / (int)blockDim.x;
        const long long base = t * PFX_TILE + (long long)threadIdx.x * per;
        T vals[16];
        T s = 0;
        for (int e = 0; e < per; ++e) {
            const long long i = base + e;
            vals[e] = i < n ? in[i] : T(0);
            s += vals[e];
        }
        T tot;
        T ex = block_exclusive<T>(s, &tot);
        T run = offset + ex;
        for (int e = 0; e < per; ++e) {
            const long long i = base + e;
            if (i < n) out[i] = run;
            run += vals[e];
        }
    }
}

extern "C" __global__ void __launch_bounds__(1024) PfxScanReduce(int *pin, int N, int *tileSums) {
    tile_reduce<unsigned>((const unsigned *)pin, N < 0 ? 0 : N, (unsigned *)tileSums);
}

extern "C" __global__ void __launch_bounds__(1024) PfxScanApply(int *pin, int N, int *tileSums, int *pout) {
    if (PFX_TILE % blockDim.x != 0 || PFX_TILE / blockDim.x > 16) return;  // geometry contract
    tile_apply<unsigned>((const unsigned *)pin, N < 0 ? 0 : N, (const unsigned *)tileSums, (unsigned *)pout);
}

extern "C" __global__ void __launch_bounds__(1024) PfxScanReduce64(long long *pin, int N, long long *tileSums) {
    tile_reduce<unsigned long long>((const unsigned long long *)pin, N < 0 ? 0 : N,
                                    (unsigned long long *)tileSums);
}

extern "C" __global__ void __launch_bounds__(1024) PfxScanApply64(long long *pin, int N, long long *tileSums,
                                                                  long long *pout) {
    if (PFX_TILE % blockDim.x != 0 || PFX_TILE / blockDim.x > 16) return;
    tile_apply<unsigned long long>((const unsigned long long *)pin, N < 0 ? 0 : N,
                                   (const unsigned long long *)tileSums, (unsigned long long *)pout);
}
