// Dense FP64 products for the SCF on small matrices (n up to a few hundred), where a library GEMM
// launches too few workgroups to use the chip:
//   dgemm_small: C = alpha op(A) op(B) + beta C (column-major, as BLAS), one block of four
//   wavefronts per 16 x 16 tile of C, splitting k, on the FP64 matrix cores (v_mfma_f64_16x16x4_f64).
// The batched forms of the SCF step (products, SP2 purification) are in batch.hip and share the tile
// code of dense_tile.h.
#include "dense_tile.h"
#include "hf_internal.h"

namespace {

// One 16 x 16 tile of C per block of 4 wavefronts, which split the k range.
template <bool TA, bool TB>
__global__ void __launch_bounds__(256) dgemm16_kernel(int m, int n, int k, double alpha, const double* __restrict__ A,
                                                      int lda, const double* __restrict__ B, int ldb, double beta,
                                                      double* __restrict__ C, int ldc) {
    __shared__ hfgpu::v4d part[3][64];
    hfgpu::gemm_tile<TA, TB>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, blockIdx.x * 16, blockIdx.y * 16, part);
}

}  // namespace

int hf_dgemm_small(hf_ctx* ctx, hipStream_t st, bool ta, bool tb, int m, int n, int k, double alpha, const double* A,
                   int lda, const double* B, int ldb, double beta, double* C, int ldc) {
    const dim3 grid((m + 15) / 16, (n + 15) / 16);
    if (ta && tb) hipLaunchKernelGGL((dgemm16_kernel<true, true>), grid, dim3(256), 0, st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    else if (ta) hipLaunchKernelGGL((dgemm16_kernel<true, false>), grid, dim3(256), 0, st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    else if (tb) hipLaunchKernelGGL((dgemm16_kernel<false, true>), grid, dim3(256), 0, st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    else hipLaunchKernelGGL((dgemm16_kernel<false, false>), grid, dim3(256), 0, st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    HF_HIP_CHECK(ctx, hipGetLastError());
    return HF_OK;
}
