// In-core J/K: the ERIs are evaluated once per SCF into HBM and every Fock build contracts them
// with the density, which is the reference's own scheme (store every unique ERI,
// HartreeFock/IntegralsRepository.cpp:566-584, then loop over the store each iteration,
// RestrictedHartreeFock::InitFockMatrix, HartreeFock/RestrictedHartreeFock.cpp:189-202, and
// UnrestrictedHartreeFock::InitFockMatrices, HartreeFock/UnrestrictedHartreeFock.cpp:228-238).
//
// Layout.  G is the symmetric matrix of ERIs over canonical function pairs,
// G[P * np + Q] = (ij|kl), P = i(i+1)/2 + j (i >= j), Q = k(k+1)/2 + l (k >= l), np = n(n+1)/2:
// twice the reference's 8-fold packed store, so that a Fock build reads every row P as one
// contiguous stream and needs no scatter.  Row P (the bra pair (i,j)) is the lower triangle of the
// symmetric n x n matrix B_P[k][l] = (ij|kl), stored row by row.
//
// Per row P, with D symmetric:
//   J[i][j] = J[j][i] = sum_{k>=l} w_kl B_P[k][l] Dtot[k][l]      (w = 2 off the diagonal, 1 on it)
//   K_d[i][:] += B_P D_d[j][:],   K_d[j][:] += B_P D_d[i][:]  (the second only when i != j)
// B_P x splits into a row part (y[k] += B[k][l] x[l], l <= k: a sum across the lanes that hold one
// row) and a column part (y[l] += B[k][l] x[k], l < k: a register of the lane that owns column l).
// One workgroup per P: four wavefronts take rows k = w (mod 4), four rows at a time; lane L owns
// columns L + 64 m.  The 8 (RHF) or 16 (UHF) row sums of four rows are reduced together by
// recursive halving (each exchange step halves the values a lane carries), 10 or 17 shuffles
// instead of 6 per value.  J_ij needs no atomics (one block per P); each block adds its two K rows
// with coalesced FP64 atomics.
//
// The kernel streams np^2 elements (8 or 4 bytes) per build: HBM-bound, ~2 + 4 ndens flops per
// element.  Each element is used once per density, so there is no operand reuse for the matrix cores.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>

#include "hf_internal.h"

namespace {

__global__ void incore_fill_kernel(double* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Dw[k(k+1)/2 + l] = w_kl * sum_d D_d[k][l]
__global__ void incore_dw_kernel(const double* __restrict__ D, int n, int ndens, double* __restrict__ Dw) {
    const int k = blockIdx.x;
    const size_t nn = (size_t)n * n;
    for (int l = threadIdx.x; l <= k; l += blockDim.x) {
        double s = 0.0;
        for (int d = 0; d < ndens; ++d) s += D[d * nn + (size_t)k * n + l];
        Dw[(size_t)k * (k + 1) / 2 + l] = (l == k ? 1.0 : 2.0) * s;
    }
}

// Sum of V[0..CNT) across the wavefront by recursive halving: at exchange step S a lane keeps the
// half of its values selected by bit S of its lane id and adds the partner's copy of that half.
// Afterwards V[0] of lane L holds the total of value index L >> (6 - log2 CNT).
template <int CNT, int S>
struct Halving {
    template <int NV>
    __device__ static __forceinline__ void run(double (&V)[NV], int lane) {
        if constexpr (CNT > 1) {
            constexpr int H = CNT / 2;
            const bool up = (lane & S) != 0;
#pragma unroll
            for (int t = 0; t < H; ++t) {
                const double keep = up ? V[H + t] : V[t];
                const double send = up ? V[t] : V[H + t];
                V[t] = keep + __shfl_xor(send, S, 64);
            }
            if constexpr (S > 1) Halving<H, S / 2>::run(V, lane);
        } else {
            V[0] += __shfl_xor(V[0], S, 64);
            if constexpr (S > 1) Halving<1, S / 2>::run(V, lane);
        }
    }
};

template <int NV>
constexpr int log2c() {
    return NV <= 1 ? 0 : 1 + log2c<NV / 2>();
}

template <int NDENS, int M, int ROWS, typename T>
__global__ void __launch_bounds__(256) incore_jk_kernel(const T* __restrict__ G, int n, int64_t np,
                                                        const double* __restrict__ D, const double* __restrict__ Dw,
                                                        double* __restrict__ J, double* __restrict__ K) {
    extern __shared__ double lds[];
    const int64_t P = blockIdx.x;
    int i = (int)((sqrt(8.0 * (double)P + 1.0) - 1.0) * 0.5);
    if ((int64_t)i * (i + 1) / 2 > P) --i;
    if ((int64_t)(i + 1) * (i + 2) / 2 <= P) ++i;
    const int j = (int)(P - (int64_t)i * (i + 1) / 2);
    double* xj = lds;               // D_d[j][:]
    double* xi = xj + NDENS * n;    // D_d[i][:]
    double* rI = xi + NDENS * n;    // additions to K_d[i][:]
    double* rJ = rI + NDENS * n;    // additions to K_d[j][:]
    double* red = rJ + NDENS * n;   // per-wavefront partials of J_ij
    const size_t nn = (size_t)n * n;
    for (int t = threadIdx.x; t < NDENS * n; t += blockDim.x) {
        const int d = t / n, c = t - d * n;
        xj[t] = D[d * nn + (size_t)j * n + c];
        xi[t] = D[d * nn + (size_t)i * n + c];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double xjl[NDENS][M], xil[NDENS][M], cI[NDENS][M], cJ[NDENS][M];
#pragma unroll
    for (int d = 0; d < NDENS; ++d)
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const int l = lane + 64 * m;
            xjl[d][m] = l < n ? xj[d * n + l] : 0.0;
            xil[d][m] = l < n ? xi[d * n + l] : 0.0;
            cI[d][m] = 0.0;
            cJ[d][m] = 0.0;
        }
    double jacc = 0.0;
    const T* __restrict__ row0 = G + P * np;
    constexpr int NV = 2 * ROWS * NDENS;           // ROWS rows x (K_i, K_j) x densities
    constexpr int SH = 6 - log2c<NV>();
    for (int k0 = w; k0 < n; k0 += 4 * ROWS) {
        double V[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) V[v] = 0.0;
        // rows k0 + 4r (r < ROWS); the rows' elements of one column chunk are loaded together
        int base[ROWS], kr[ROWS];                        // row offsets fit 32 bits: np < 2^31 for n <= 512
        double xkj[ROWS][NDENS], xki[ROWS][NDENS];
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
            const int k = k0 + 4 * r;
            kr[r] = k < n ? k : -1;                // -1: past the last row (no element)
            const int kk = k < n ? k : n - 1;
            base[r] = kk * (kk + 1) / 2;
#pragma unroll
            for (int d = 0; d < NDENS; ++d) {
                xkj[r][d] = xj[d * n + kk];
                xki[r][d] = xi[d * n + kk];
            }
        }
        const int klast = k0 + 4 * (ROWS - 1) < n ? k0 + 4 * (ROWS - 1) : n - 1;
#pragma unroll
        for (int m = 0; m < M; ++m) {
            if (64 * m > klast) break;             // uniform across the wavefront
            const int l = lane + 64 * m;
            double gv[ROWS], dv[ROWS];
#pragma unroll
            for (int r = 0; r < ROWS; ++r) {
                const int kk = kr[r] < 0 ? n - 1 : kr[r];
                const int at = base[r] + (l <= kk ? l : kk);     // clamped: always inside the row
                // 32-bit byte offsets from the block-uniform row and density bases (address arithmetic
                // stays in one VALU op; a row is at most np * 8 < 2^20 bytes)
                const double gx = (double)*(const T*)((const char*)row0 + (unsigned)at * (unsigned)sizeof(T));
                dv[r] = *(const double*)((const char*)Dw + (unsigned)at * 8u);
                gv[r] = l <= kr[r] ? gx : 0.0;
            }
#pragma unroll
            for (int r = 0; r < ROWS; ++r) {
                const double g = gv[r];
                jacc = fma(g, dv[r], jacc);
#pragma unroll
                for (int d = 0; d < NDENS; ++d) {
                    V[(2 * r) * NDENS + d] = fma(g, xjl[d][m], V[(2 * r) * NDENS + d]);
                    V[(2 * r + 1) * NDENS + d] = fma(g, xil[d][m], V[(2 * r + 1) * NDENS + d]);
                }
                const double gc = l < kr[r] ? g : 0.0;            // the diagonal element has no column part
#pragma unroll
                for (int d = 0; d < NDENS; ++d) {
                    cI[d][m] = fma(gc, xkj[r][d], cI[d][m]);
                    cJ[d][m] = fma(gc, xki[r][d], cJ[d][m]);
                }
            }
        }
        Halving<NV, 32>::run(V, lane);
        if ((lane & ((1 << SH) - 1)) == 0) {
            const int idx = lane >> SH;
            const int r = idx / (2 * NDENS), a = (idx / NDENS) & 1, d = idx % NDENS;
            const int k = k0 + 4 * r;
            if (k < n) (a ? rJ : rI)[d * n + k] = V[0];   // row k belongs to this wavefront alone
        }
    }
    // J_ij: wavefront sums, then the block's
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) jacc += __shfl_xor(jacc, s, 64);
    if (lane == 0) red[w] = jacc;
    __syncthreads();
    // column parts of every wavefront
#pragma unroll
    for (int d = 0; d < NDENS; ++d)
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const int l = lane + 64 * m;
            if (l < n) {
                atomicAdd(&rI[d * n + l], cI[d][m]);
                atomicAdd(&rJ[d * n + l], cJ[d][m]);
            }
        }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double v = (red[0] + red[1]) + (red[2] + red[3]);
        J[(size_t)i * n + j] += v;
        if (i != j) J[(size_t)j * n + i] += v;
    }
    for (int t = threadIdx.x; t < NDENS * n; t += blockDim.x) {
        const int d = t / n, c = t - d * n;
        atomicAdd(&K[d * nn + (size_t)i * n + c], rI[t]);
        if (i != j) atomicAdd(&K[d * nn + (size_t)j * n + c], rJ[t]);
    }
}

template <int NDENS, int ROWS, typename T>
hipError_t launch_incore(const T* G, int n, int64_t np, const double* D, const double* Dw, double* J, double* K,
                         hipStream_t st) {
    const size_t lds = (4 * (size_t)NDENS * n + 4) * sizeof(double);
    const dim3 grid((unsigned)np), block(256);
    switch ((n + 63) / 64) {
#define HF_INCORE_CASE(MM)                                                                                          \
    case MM:                                                                                                        \
        hipLaunchKernelGGL((incore_jk_kernel<NDENS, MM, ROWS, T>), grid, block, lds, st, G, n, np, D, Dw, J, K); \
        break;
        HF_INCORE_CASE(1)
        HF_INCORE_CASE(2)
        HF_INCORE_CASE(3)
        HF_INCORE_CASE(4)
        HF_INCORE_CASE(5)
        HF_INCORE_CASE(6)
        HF_INCORE_CASE(7)
        HF_INCORE_CASE(8)
#undef HF_INCORE_CASE
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Row-group height of the contraction (rows of one wavefront whose loads are issued together and
// whose row sums are reduced together): 2, 4 or 8.
template <typename T>
hipError_t launch_incore_rows(int rows, int ndens, const T* G, int n, int64_t np, const double* D, const double* Dw,
                              double* J, double* K, hipStream_t st) {
    if (ndens == 1) {
        if (rows == 2) return launch_incore<1, 2>(G, n, np, D, Dw, J, K, st);
        if (rows == 8) return launch_incore<1, 8>(G, n, np, D, Dw, J, K, st);
        return launch_incore<1, 4>(G, n, np, D, Dw, J, K, st);
    }
    if (rows == 2) return launch_incore<2, 2>(G, n, np, D, Dw, J, K, st);
    if (rows == 8) return launch_incore<2, 8>(G, n, np, D, Dw, J, K, st);
    return launch_incore<2, 4>(G, n, np, D, Dw, J, K, st);
}

// Autotuner: the row-group height is picked per (column chunks M = ceil(n/64), densities, store
// precision) by timing every candidate on the caller's own store, density and stream, once per
// process (the table is shared by all contexts; a server's slots tune each key once).  Candidates
// write to scratch J/K, so the build's outputs are untouched.
struct IncoreTuning {
    std::mutex mu;
    std::map<int, int> rows;                  // key -> chosen rows
};
IncoreTuning& tuning() {
    static IncoreTuning t;
    return t;
}

template <typename T>
int tuned_rows(hf_ctx*, int ndens, const T* G, int n, int64_t np, const double* D, const double* Dw,
               hipStream_t st) {
    const int key = ((n + 63) / 64) * 100 + ndens * 10 + (sizeof(T) == 4 ? 1 : 0);
    {
        std::lock_guard<std::mutex> lock(tuning().mu);
        auto it = tuning().rows.find(key);
        if (it != tuning().rows.end()) return it->second;
    }
    const size_t nn = (size_t)n * n;
    DevBuf<double> sJ, sK;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int best = 4;
    float best_ms = 0;
    bool ok = sJ.alloc(nn) == hipSuccess && sK.alloc(2 * nn) == hipSuccess &&
              hipEventCreate(&ev[0]) == hipSuccess && hipEventCreate(&ev[1]) == hipSuccess;
    for (int rows : {2, 4, 8}) {
        if (!ok) break;
        float t = 0;
        for (int rep = 0; rep < 3 && ok; ++rep) {          // first run warms the caches and the code
            ok = hipEventRecord(ev[0], st) == hipSuccess &&
                 launch_incore_rows(rows, ndens, G, n, np, D, Dw, sJ.p, sK.p, st) == hipSuccess &&
                 hipEventRecord(ev[1], st) == hipSuccess && hipEventSynchronize(ev[1]) == hipSuccess;
            float ms = 0;
            if (ok) ok = hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess;
            if (rep > 0) t = rep == 1 ? ms : std::min(t, ms);
        }
        if (ok && (best_ms == 0 || t < best_ms)) {
            best_ms = t;
            best = rows;
        }
    }
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipGetLastError();
    if (!ok) return 4;                                     // tuning failed: the default, not recorded
    std::lock_guard<std::mutex> lock(tuning().mu);
    tuning().rows[key] = best;
    return best;
}

}  // namespace

void hf_release_incore(hf_ctx* ctx, size_t keep_bytes) {
    if (ctx->d_G.n * sizeof(double) + ctx->d_G32.n * sizeof(float) > keep_bytes) {
        ctx->d_G.release();
        ctx->d_G32.release();
    }
}

int hf_incore_jk_impl(hf_ctx* ctx, int ndens, const double* G, const float* g32, const double* d_D, double* d_J,
                      double* d_K, hipStream_t st, bool accumulate) {
    if (!ctx->have_mol) return ctx->fail(HF_E_STATE, "no molecule set");
    if (ndens < 1 || ndens > 2) return ctx->fail(HF_E_ARG, "ndens must be 1 or 2");
    const int n = ctx->mol.nbf;
    if (n > kIncoreMaxN) return ctx->fail(HF_E_UNSUPPORTED, "in-core J/K takes at most 512 basis functions");
    const int64_t np = (int64_t)n * (n + 1) / 2;
    const size_t nn = (size_t)n * n;
    HF_HIP_CHECK(ctx, ctx->d_Dw.alloc((size_t)np));
    hipLaunchKernelGGL(incore_dw_kernel, dim3(n), dim3(256), 0, st, d_D, n, ndens, ctx->d_Dw.p);
    if (!accumulate) {
        hipLaunchKernelGGL(incore_fill_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, d_J, nn);
        hipLaunchKernelGGL(incore_fill_kernel, dim3((unsigned)((ndens * nn + 255) / 256)), dim3(256), 0, st, d_K,
                           ndens * nn);
    }
    // row-group height from the autotuner (first build of each kind in the process times the candidates)
    const int rows = g32 ? tuned_rows(ctx, ndens, g32, n, np, d_D, ctx->d_Dw.p, st)
                         : tuned_rows(ctx, ndens, G, n, np, d_D, ctx->d_Dw.p, st);
    // statistics, and with profile = 1 events around the contraction (read at the next flush)
    const int f = g32 ? 1 : 0;
    const bool prof = ctx->opt.profile != 0;
    if (prof) {
        if (ctx->incore_pending >= 0) {
            HF_HIP_CHECK(ctx, hipEventSynchronize(ctx->incore_ev[1]));
            float ms = 0;
            HF_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->incore_ev[0], ctx->incore_ev[1]));
            ctx->incore_stat[ctx->incore_pending].ms += ms;
            ctx->incore_pending = -1;
        }
        for (hipEvent_t& ev : ctx->incore_ev)
            if (!ev) HF_HIP_CHECK(ctx, hipEventCreate(&ev));
        HF_HIP_CHECK(ctx, hipEventRecord(ctx->incore_ev[0], st));
    }
    KernelStat& ks = ctx->incore_stat[f];
    ks.launches += 1;
    ks.eris += (double)np * (double)np;                                  // elements streamed
    ks.flops += (double)np * (double)np * (2.0 + 8.0 * ndens);           // J, and row + column parts of K
    hipError_t e;
    if (g32)
        e = launch_incore_rows(rows, ndens, g32, n, np, d_D, ctx->d_Dw.p, d_J, d_K, st);
    else
        e = launch_incore_rows(rows, ndens, G, n, np, d_D, ctx->d_Dw.p, d_J, d_K, st);
    if (e != hipSuccess) return ctx->fail(HF_E_HIP, std::string("in-core J/K: ") + hipGetErrorString(e));
    if (prof) {
        HF_HIP_CHECK(ctx, hipEventRecord(ctx->incore_ev[1], st));
        ctx->incore_pending = f;
    }
    return HF_OK;
}
