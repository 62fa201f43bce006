// vector.hip -- vector task kernels for iterated graphs (code object vector.hsaco).
//
//   ScaleInPlace(float* A, float s, int n)   A[i] = s * A[i] for i < n
//
// The in-place vector task of ptaskUnitTest/iteration.cpp's simple-iteration graph (an in/out
// vector port, a sticky scalar, the element count from a port derived from the vector).  Any
// launch geometry: threads stride over the elements.
#include <hip/hip_runtime.h>

extern "C" __global__ void __launch_bounds__(1024) ScaleInPlace(float *A, float s, int n) {
    const int stride = (int)(gridDim.x * blockDim.x);
    for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += stride) A[i] = s * A[i];
}
