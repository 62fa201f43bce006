// gmm.hip -- the EM-GMM task kernels (code object gmm.hsaco); the bodies are in gmm_body.h.
//
//   GmmAccumulate(const float* X, GMM_DIMS d, const double* model, double* partials)
//     grid d.blocks blocks of 256; block b accumulates the tiles of unit b into partial row b.
//   GmmAccumulateMfma(...)   the same task with its products on the matrix cores (gmm_mfma.h)
//   GmmFinalize(double* model, const double* partials, GMM_DIMS d)
//     grid K + 1 blocks of 256; block k < K updates component k, block K the log-likelihood.
#include "gmm_body.h"
#include "gmm_mfma.h"

extern "C" __global__ void __launch_bounds__(256)
GmmAccumulate(const float *__restrict__ X, GMM_DIMS d, const double *__restrict__ model,
              double *__restrict__ partials) {
    __shared__ __align__(16) GmmShared sh;
    if (!gmm_dims_ok(d) || blockDim.x != 256 || (int)gridDim.x != d.blocks) return;
    gmm_accumulate_unit(X, d, model, partials, (int)blockIdx.x, sh);
}

extern "C" __global__ void __launch_bounds__(256, 2)
GmmAccumulateMfma(const float *__restrict__ X, GMM_DIMS d, const double *__restrict__ model,
                  double *__restrict__ partials) {
    __shared__ __align__(16) GmmMfmaShared sh;
    if (!gmm_dims_ok(d) || blockDim.x != 256 || (int)gridDim.x != d.blocks) return;
    gmm_mfma_unit(X, d, model, partials, (int)blockIdx.x, sh);
}

extern "C" __global__ void __launch_bounds__(256)
GmmFinalize(double *__restrict__ model, const double *__restrict__ partials, GMM_DIMS d) {
    if (!gmm_dims_ok(d) || blockDim.x != 256 || (int)gridDim.x != d.K + 1) return;
    gmm_finalize_block(model, partials, d, (int)blockIdx.x);
}
