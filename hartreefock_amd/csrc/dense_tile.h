// 16 x 16 FP64 tile products on the matrix cores, shared by the one-matrix kernels (dense.hip) and the
// batched ones (batch.hip), so a product computed in a batch is bit-identical to the same product alone.
//
// Matrix-core lane layout of __builtin_amdgcn_mfma_f64_16x16x4f64 (A 16x4, B 4x16, D 16x16):
//   lane l holds A[l % 16][l / 16] and B[l / 16][l % 16]; element r of its 4-vector result is
//   D[l / 16 + 4 r][l % 16].
#pragma once
#include <hip/hip_runtime.h>

namespace hfgpu {

typedef double v4d __attribute__((ext_vector_type(4)));

// Partial product of a 16 x 16 tile over k in [kb, ke) for the calling wavefront: fa(i, k) and
// fb(k, j) return elements of op(A) and op(B) (zero outside).  8 products per loop trip, with their
// 16 loads issued first.
template <class FA, class FB>
__device__ __forceinline__ v4d tile_product(int kb, int ke, int i0, int j0, FA fa, FB fb) {
    const int l = threadIdx.x & 63, lr = l & 15, lk = l >> 4;
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = kb; k0 < ke; k0 += 32) {
        double a[8], b[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int kk = k0 + 4 * u + lk;
            a[u] = kk < ke ? fa(i0 + lr, kk) : 0.0;
            b[u] = kk < ke ? fb(kk, j0 + lr) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u], acc, 0, 0, 0);
    }
    return acc;
}

// Sum of the four wavefronts' partial tiles, returned to wavefront 0 (others get garbage).
__device__ __forceinline__ v4d reduce_waves(v4d acc, v4d (*part)[64]) {
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (w > 0) part[w - 1][l] = acc;
    __syncthreads();
    if (w == 0)
#pragma unroll
        for (int q = 0; q < 3; ++q) acc += part[q][l];
    return acc;
}

// k range of wavefront w of 4: contiguous strips of a multiple of 4.
__device__ __forceinline__ void k_strip(int k, int w, int& kb, int& ke) {
    const int q = ((k + 15) / 16) * 4;
    kb = min(k, w * q);
    ke = min(k, kb + q);
}

// Tile (i0, j0) of C = alpha op(A) op(B) + beta C (column-major) by the block's four wavefronts, which split
// the k range; wavefront 0 writes the tile.
template <bool TA, bool TB>
__device__ __forceinline__ void gemm_tile(int m, int n, int k, double alpha, const double* __restrict__ A, int lda,
                                          const double* __restrict__ B, int ldb, double beta, double* __restrict__ C,
                                          int ldc, int i0, int j0, v4d (*part)[64]) {
    const int l = threadIdx.x & 63, lr = l & 15, lk = l >> 4, w = threadIdx.x >> 6;
    int kb, ke;
    k_strip(k, w, kb, ke);
    auto fa = [&](int i, int kk) -> double {
        return i < m ? (TA ? A[kk + (size_t)i * lda] : A[i + (size_t)kk * lda]) : 0.0;
    };
    auto fb = [&](int kk, int j) -> double {
        return j < n ? (TB ? B[j + (size_t)kk * ldb] : B[kk + (size_t)j * ldb]) : 0.0;
    };
    v4d acc = reduce_waves(tile_product(kb, ke, i0, j0, fa, fb), part);
    if (w != 0) return;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = i0 + lk + 4 * r, j = j0 + lr;
        if (i < m && j < n) {
            double* c = C + i + (size_t)j * ldc;
            *c = beta == 0.0 ? alpha * acc[r] : alpha * acc[r] + beta * *c;
        }
    }
}

}  // namespace hfgpu
