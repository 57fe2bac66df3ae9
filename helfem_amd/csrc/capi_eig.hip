// C-ABI batched symmetric eigensolver (include/hfgpu.h): the per-block
// generalised eigenproblems of the SCF drivers, reduced to standard form by
// the orthogonaliser, solved for a strided batch at once with rocSOLVER's
// divide-and-conquer dsyevd on the context stream.  The reference solves each
// block on the host with Eigen's SelfAdjointEigenSolver on the orthogonal
// basis F' = Sinvh' F Sinvh (eig_gsym / eig_gsym_sub, src/general/
// scf_helpers.cpp:41-80).
#include "capi_common.hpp"
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

using namespace hfg;

namespace {
rocblas_handle blas_of(hf_ctx *ctx) {
  if (!ctx->blas) {
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) throw Error(HF_EDEVICE, "rocblas_create_handle failed");
    ctx->blas = h;
  }
  rocblas_handle h = (rocblas_handle)ctx->blas;
  if (rocblas_set_stream(h, ctx->cur) != rocblas_status_success) throw Error(HF_EDEVICE, "rocblas_set_stream failed");
  return h;
}
}  // namespace

void hf_ctx_release_blas(hf_ctx *ctx) {
  if (ctx && ctx->blas) {
    (void)rocblas_destroy_handle((rocblas_handle)ctx->blas);
    ctx->blas = nullptr;
  }
}

extern "C" {

int hf_eigh_batched(hf_ctx *ctx, int n, double *dA, int64_t lda, int64_t strideA, double *dW, double *dE,
                    int *dinfo, int batch) {
  return guarded([&] {
    need(ctx != nullptr, "ctx is NULL");
    need(n >= 1 && batch >= 1, "n and batch must be positive");
    need(lda >= n && strideA >= lda * (int64_t)n, "lda < n or the batch stride overlaps");
    need(dA != nullptr && dW != nullptr && dE != nullptr && dinfo != nullptr, "NULL device pointer");
    DeviceGuard g(ctx->device);
    const rocblas_status st = rocsolver_dsyevd_strided_batched(blas_of(ctx), rocblas_evect_original, rocblas_fill_lower,
                                                               n, dA, (rocblas_int)lda, strideA, dW, n, dE, n, dinfo,
                                                               batch);
    if (st != rocblas_status_success) throw Error(HF_EDEVICE, "rocsolver_dsyevd_strided_batched failed");
  });
}

}  // extern "C"
