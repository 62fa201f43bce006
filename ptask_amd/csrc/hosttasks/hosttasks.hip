// hosttasks.hip -- libptk_hosttasks.so: host tasks with a dependent HIP accelerator, the roles of
// the reference's HostTasks library (HostTasks/HostTasks.cpp:238-400) on this build.
//
//   SGemm(LPDEPENDENTCONTEXT)   EX form: C = A B (column-major, SGEMMPARMS {ax, ay, bx, by, cx, cy}:
//                               m = ax, n = by, k = bx, lda = ax, ldb = bx, ldc = cx) by rocBLAS on
//                               the dependent stream; arguments A, B, C(in/out), parms
//                               (HostTasks.cpp:298-337 does the same with cuBLAS)
//   ScaleDeviceVector(nArgs, args, isDep, devBindings, nDeps, deps)   legacy form: A[i] *= s for
//                               i < n on the device, A a device pointer; the call returns when the
//                               work is done
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdint>
#include <cstring>

#include "hosttask.h"

typedef struct _sgemm_parms_t {
    int ax, ay, bx, by, cx, cy;
} SGEMMPARMS;

extern "C" void SGemm(LPDEPENDENTCONTEXT ctx) {
    if (!ctx || ctx->nArguments != 4 || !ctx->pbIsDependentBinding[0] || !ctx->pbIsDependentBinding[1] ||
        !ctx->pbIsDependentBinding[2])
        return;  // this build has no host fallback for SGemm
    const SGEMMPARMS *p = (const SGEMMPARMS *)ctx->ppArguments[3];
    const float *A = (const float *)ctx->ppArguments[0], *B = (const float *)ctx->ppArguments[1];
    float *C = (float *)ctx->ppArguments[2];
    const float alpha = 1.0f, beta = 0.0f;
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) return;
    rocblas_set_stream(h, (hipStream_t)ctx->pStreams[0]);
    rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_none, p->ax, p->by, p->bx, &alpha, A, p->ax, B, p->bx,
                  &beta, C, p->cx);
    rocblas_destroy_handle(h);  // the work stays queued on the stream
}

__global__ void ScaleKernel(float *A, float s, int n) {
    for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += (int)(gridDim.x * blockDim.x)) A[i] *= s;
}

extern "C" void ScaleDeviceVector(unsigned int nArgs, void **args, BOOL *isDep, void **devBindings, unsigned int nDeps,
                                  void **deps) {
    (void)devBindings;
    (void)deps;
    if (nArgs != 3 || nDeps != 1 || !isDep[0]) return;
    float s;
    const uintptr_t bits = (uintptr_t)args[1];
    memcpy(&s, &bits, sizeof(s));
    const int n = (int)(intptr_t)args[2];
    if (n <= 0) return;
    const unsigned grid = (unsigned)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
    hipLaunchKernelGGL(ScaleKernel, dim3(grid), dim3(256), 0, 0, (float *)args[0], s, n);
    (void)hipStreamSynchronize(0);
}
