// ptk_util.hip -- small utility kernels of the HIP backend (code object ptk_util.hsaco).
//
// PtkCopyWords streams a buffer to another with 16-byte accesses.  It is the runtime's
// device-to-device block copy for datablock views and the reference point for HBM traffic
// counters: a copy of B bytes moves exactly B bytes in and B bytes out.
#include <hip/hip_runtime.h>

extern "C" __global__ void __launch_bounds__(256)
PtkCopyWords(float4 *__restrict__ dst, const float4 *__restrict__ src, unsigned long long n16) {
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
         i += stride)
        dst[i] = src[i];
}

// Fills n 4-byte words with `value` (used instead of a memset node inside captured graphs).
extern "C" __global__ void __launch_bounds__(256)
PtkFillWords(unsigned *__restrict__ dst, unsigned long long n, unsigned value) {
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += stride)
        dst[i] = value;
}

// Streaming probes of the HBM rate for other access mixes (ptk_copy --mode): PtkStreamMix reads
// six arrays and writes three (the half-step FDTD mix, without a stencil); PtkReadWords only
// reads (its store never happens for finite data).
extern "C" __global__ void __launch_bounds__(256)
PtkStreamMix(const float4 *__restrict__ a0, const float4 *__restrict__ a1, const float4 *__restrict__ a2,
             const float4 *__restrict__ a3, const float4 *__restrict__ a4, const float4 *__restrict__ a5,
             float4 *__restrict__ o0, float4 *__restrict__ o1, float4 *__restrict__ o2, unsigned long long n16) {
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) {
        const float4 x0 = a0[i], x1 = a1[i], x2 = a2[i], x3 = a3[i], x4 = a4[i], x5 = a5[i];
        o0[i] = make_float4(x0.x + x3.x, x0.y + x3.y, x0.z + x3.z, x0.w + x3.w);
        o1[i] = make_float4(x1.x + x4.x, x1.y + x4.y, x1.z + x4.z, x1.w + x4.w);
        o2[i] = make_float4(x2.x + x5.x, x2.y + x5.y, x2.z + x5.z, x2.w + x5.w);
    }
}

extern "C" __global__ void __launch_bounds__(256)
PtkReadWords(const float4 *__restrict__ src, unsigned long long n16, float *__restrict__ sink) {
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    float acc = 0.f;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) {
        const float4 v = src[i];
        acc += v.x + v.y + v.z + v.w;
    }
    if (acc != acc) sink[0] = acc;  // NaN only: keeps the loads without a store
}
