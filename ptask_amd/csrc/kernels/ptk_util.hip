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
