// In-core J/K: the ERIs are evaluated once per SCF into HBM and every Fock build contracts them
// with the density, which is the reference's own scheme (store every unique ERI,
// HartreeFock/IntegralsRepository.cpp:566-584, then loop over the store each iteration,
// RestrictedHartreeFock::InitFockMatrix, HartreeFock/RestrictedHartreeFock.cpp:189-202, and
// UnrestrictedHartreeFock::InitFockMatrices, HartreeFock/UnrestrictedHartreeFock.cpp:228-238).
//
// Layout.  G is the symmetric matrix of ERIs over canonical function pairs: G[P * np + Q] = (ij|kl)
// with P = pair_pad_index(i, j), Q = pair_pad_index(k, l) (i >= j, k >= l): the pair triangle row by
// row, each row padded to an even length, np = pair_pad_base(n) ~ n(n+1)/2 + n/2.  Each unique ERI
// is held twice (twice the reference's 8-fold packed store), so a Fock build reads every row P as one
// contiguous stream and needs no scatter; the padding makes every even column offset 16-byte aligned.
// Row P (the bra pair (i,j)) is the lower triangle of the symmetric n x n matrix B_P[k][l] = (ij|kl).
//
// Per row P, with D symmetric:
//   J[i][j] = J[j][i] = sum_{k>=l} w_kl B_P[k][l] Dtot[k][l]      (w = 2 off the diagonal, 1 on it)
//   K_d[i][:] += B_P D_d[j][:],   K_d[j][:] += B_P D_d[i][:]  (the second only when i != j)
// B_P x splits into a row part (y[k] += B[k][l] x[l], l <= k: a sum across the lanes that hold one
// row) and a column part (y[l] += B[k][l] x[k], l < k: a register of the lane that owns column l).
// One workgroup per P: four wavefronts take rows k = w (mod 4), ROWS rows at a time; lane L owns the
// column pairs (2L, 2L+1) + 128 m and reads each pair with one load; the loads of every column chunk
// of a group go out together (optionally the next group's too, before this group's arithmetic).
// The 2 ROWS ndens row sums of a group are reduced together by recursive halving (each exchange step
// halves the values a lane carries), 10 exchanges for 8 values instead of 6 per value, and every
// exchange stays in registers (gfx950 permlane swaps and DPP moves, no LDS round trip).  J_ij needs no
// atomics (one block per P); each block adds its two K rows with coalesced FP64 atomics.
//
// The kernel streams np^2 elements (8 or 4 bytes) per build: HBM-bound, ~2 + 8 ndens flops per
// element.  Each element is used once per density, so there is no operand reuse for the matrix cores.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>

#include "hf_internal.h"
#include "wave_reduce.h"

namespace {

using hfgpu::Halving;
using hfgpu::log2c;

__global__ void incore_fill_kernel(double* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Dw[pair_pad_index(k, l)] = w_kl * sum_d D_d[k][l] (the store's padded row order; padding is 0)
__global__ void incore_dw_kernel(const double* __restrict__ D, int n, int ndens, double* __restrict__ Dw) {
    const int k = blockIdx.x;
    const size_t nn = (size_t)n * n;
    const int64_t b = hfgpu::pair_pad_base(k), len = hfgpu::pair_pad_base(k + 1) - b;
    for (int l = threadIdx.x; l < len; l += blockDim.x) {
        double s = 0.0;
        if (l <= k)
            for (int d = 0; d < ndens; ++d) s += D[d * nn + (size_t)k * n + l];
        Dw[b + l] = (l == k ? 1.0 : 2.0) * s;
    }
}

template <typename T>
struct PairOf;
template <>
struct PairOf<double> {
    using V = double2;
};
template <>
struct PairOf<float> {
    using V = float2;
};

// One workgroup per padded row index R of the store (R = pair_pad_index(i, j); padding rows return at
// once).  M2 = ceil(n / 128) column chunks: lane L owns columns 2L + 128 m and 2L + 1 + 128 m and reads
// both with one aligned 16-byte (FP64) or 8-byte (FP32) load.
template <int NDENS, int M2, int ROWS, bool PREFETCH, typename T>
__global__ void __launch_bounds__(256) incore_jk_kernel(const T* __restrict__ G, int n, int64_t npad,
                                                        const double* __restrict__ D, const double* __restrict__ Dw,
                                                        double* __restrict__ J, double* __restrict__ K) {
    extern __shared__ double lds[];
    const int64_t R = blockIdx.x;
    int i = (int)sqrt(2.0 * (double)R) - 1;
    if (i < 0) i = 0;
    while (hfgpu::pair_pad_base(i + 1) <= R) ++i;
    while (i > 0 && hfgpu::pair_pad_base(i) > R) --i;
    const int j = (int)(R - hfgpu::pair_pad_base(i));
    if (j > i) return;                                      // padding row: all zero (whole block)
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
    double xjl[NDENS][M2][2], xil[NDENS][M2][2], cI[NDENS][M2][2], cJ[NDENS][M2][2];
#pragma unroll
    for (int d = 0; d < NDENS; ++d)
#pragma unroll
        for (int m = 0; m < M2; ++m)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int l = 2 * lane + 128 * m + h;
                xjl[d][m][h] = l < n ? xj[d * n + l] : 0.0;
                xil[d][m][h] = l < n ? xi[d * n + l] : 0.0;
                cI[d][m][h] = 0.0;
                cJ[d][m][h] = 0.0;
            }
    double jacc = 0.0;
    // byte offsets from the block-uniform row base (a padded row is at most ~1 MB)
    const char* __restrict__ row0 = (const char*)(G + R * npad);
    const char* __restrict__ dwb = (const char*)Dw;
    constexpr int NV = 2 * ROWS * NDENS;           // ROWS rows x (K_i, K_j) x densities
    constexpr int SH = 6 - log2c<NV>();
    using PV = typename PairOf<T>::V;
    // The loads of a row group: every column chunk of its ROWS rows k0 + 4r, issued together.  Addresses
    // are clamped into the row (rows past the last one read row n - 1), so every load is in bounds;
    // elements outside the row are masked where they are used.
    auto load = [&](int k0, PV (&g)[ROWS][M2], double2 (&dd)[ROWS][M2]) {
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
            const int k = k0 + 4 * r;
            const int kk = k < n ? k : n - 1;
            const int b = (int)hfgpu::pair_pad_base(kk);
            const int kev = kk & ~1;               // last even column of the row: (kev, kev + 1) is in the row
#pragma unroll
            for (int m = 0; m < M2; ++m) {
                const int l0 = 2 * lane + 128 * m;
                const unsigned at = (unsigned)(b + (l0 <= kev ? l0 : kev));
                g[r][m] = *(const PV*)(row0 + at * (unsigned)sizeof(T));
                dd[r][m] = *(const double2*)(dwb + at * 8u);
            }
        }
    };
    PV ga[ROWS][M2];
    double2 da[ROWS][M2];
    load(w, ga, da);
    for (int k0 = w; k0 < n; k0 += 4 * ROWS) {
        // PREFETCH: the next group's loads go out before this group's arithmetic and reduction
        PV gb[ROWS][M2];
        double2 db[ROWS][M2];
        if constexpr (PREFETCH) load(k0 + 4 * ROWS, gb, db);
        double V[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) V[v] = 0.0;
        int kr[ROWS];
        double xkj[ROWS][NDENS], xki[ROWS][NDENS];
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
            const int k = k0 + 4 * r;
            kr[r] = k < n ? k : -1;                // -1: past the last row (no element)
            const int kk = k < n ? k : n - 1;
#pragma unroll
            for (int d = 0; d < NDENS; ++d) {
                xkj[r][d] = xj[d * n + kk];
                xki[r][d] = xi[d * n + kk];
            }
        }
        const int klast = k0 + 4 * (ROWS - 1) < n ? k0 + 4 * (ROWS - 1) : n - 1;
#pragma unroll
        for (int m = 0; m < M2; ++m) {
            if (128 * m > klast) break;            // uniform across the wavefront
            const int l0 = 2 * lane + 128 * m;
#pragma unroll
            for (int r = 0; r < ROWS; ++r)
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const double graw = h ? (double)ga[r][m].y : (double)ga[r][m].x;
                    const double g = l0 + h <= kr[r] ? graw : 0.0;
                    jacc = fma(g, h ? da[r][m].y : da[r][m].x, jacc);
#pragma unroll
                    for (int d = 0; d < NDENS; ++d) {
                        V[(2 * r) * NDENS + d] = fma(g, xjl[d][m][h], V[(2 * r) * NDENS + d]);
                        V[(2 * r + 1) * NDENS + d] = fma(g, xil[d][m][h], V[(2 * r + 1) * NDENS + d]);
                    }
                    const double gc = l0 + h < kr[r] ? g : 0.0;    // the diagonal element has no column part
#pragma unroll
                    for (int d = 0; d < NDENS; ++d) {
                        cI[d][m][h] = fma(gc, xkj[r][d], cI[d][m][h]);
                        cJ[d][m][h] = fma(gc, xki[r][d], cJ[d][m][h]);
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
        if constexpr (PREFETCH) {
#pragma unroll
            for (int r = 0; r < ROWS; ++r)
#pragma unroll
                for (int m = 0; m < M2; ++m) {
                    ga[r][m] = gb[r][m];
                    da[r][m] = db[r][m];
                }
        } else if (k0 + 4 * ROWS < n) {
            load(k0 + 4 * ROWS, ga, da);
        }
    }
    // J_ij: wavefront sums, then the block's
    {
        double jv[1] = {jacc};
        Halving<1, 32>::run(jv, lane);
        jacc = jv[0];
    }
    if (lane == 0) red[w] = jacc;
    __syncthreads();
    // column parts of every wavefront
#pragma unroll
    for (int d = 0; d < NDENS; ++d)
#pragma unroll
        for (int m = 0; m < M2; ++m)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int l = 2 * lane + 128 * m + h;
                if (l < n) {
                    atomicAdd(&rI[d * n + l], cI[d][m][h]);
                    atomicAdd(&rJ[d * n + l], cJ[d][m][h]);
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

template <int NDENS, int ROWS, bool PREFETCH, typename T>
hipError_t launch_incore(const T* G, int n, int64_t npad, const double* D, const double* Dw, double* J, double* K,
                         hipStream_t st) {
    const size_t lds = (4 * (size_t)NDENS * n + 4) * sizeof(double);
    const dim3 grid((unsigned)npad), block(256);
    switch ((n + 127) / 128) {
#define HF_INCORE_CASE(MM)                                                                                        \
    case MM:                                                                                                      \
        hipLaunchKernelGGL((incore_jk_kernel<NDENS, MM, ROWS, PREFETCH, T>), grid, block, lds, st, G, n, npad, D, \
                           Dw, J, K);                                                                             \
        break;
        HF_INCORE_CASE(1)
        HF_INCORE_CASE(2)
        HF_INCORE_CASE(3)
        HF_INCORE_CASE(4)
#undef HF_INCORE_CASE
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Tile shapes of the contraction: rows of one wavefront whose loads are issued together and whose
// row sums are reduced together, and whether the next group's loads are issued before the current
// group's arithmetic (more bytes in flight per wavefront for more registers, so fewer wavefronts).
enum IncoreShape { kRows2 = 0, kRows4 = 1, kRows2Prefetch = 2, kIncoreShapes = 3 };

template <int NDENS, typename T>
hipError_t launch_incore_nd(int shape, const T* G, int n, int64_t np, const double* D, const double* Dw, double* J,
                            double* K, hipStream_t st) {
    if (shape == kRows4) return launch_incore<NDENS, 4, false>(G, n, np, D, Dw, J, K, st);
    if (shape == kRows2Prefetch) return launch_incore<NDENS, 2, true>(G, n, np, D, Dw, J, K, st);
    return launch_incore<NDENS, 2, false>(G, n, np, D, Dw, J, K, st);
}

template <typename T>
hipError_t launch_incore_shape(int shape, int ndens, const T* G, int n, int64_t np, const double* D,
                               const double* Dw, double* J, double* K, hipStream_t st) {
    return ndens == 1 ? launch_incore_nd<1>(shape, G, n, np, D, Dw, J, K, st)
                      : launch_incore_nd<2>(shape, G, n, np, D, Dw, J, K, st);
}

// Autotuner: the tile shape is picked per (column chunks M2 = ceil(n/128), densities, store
// precision) by timing every candidate on the caller's own store, density and stream, once per
// process (the table is shared by all contexts; a server's slots tune each key once).  Candidates
// write to scratch J/K, so the build's outputs are untouched.
struct IncoreTuning {
    std::mutex mu;
    std::map<int, int> shape;                 // key -> chosen IncoreShape
};
IncoreTuning& tuning() {
    static IncoreTuning t;
    return t;
}

int tuning_key(int n, int ndens, bool fp32) { return ((n + 127) / 128) * 100 + ndens * 10 + (fp32 ? 1 : 0); }

// The shape the autotuner picked for this kind of build, or the default when it has not run (no timing here).
int known_shape(int n, int ndens, bool fp32) {
    std::lock_guard<std::mutex> lock(tuning().mu);
    auto it = tuning().shape.find(tuning_key(n, ndens, fp32));
    return it == tuning().shape.end() ? (int)kRows2 : it->second;
}

template <typename T>
int tuned_shape(hf_ctx*, int ndens, const T* G, int n, int64_t np, const double* D, const double* Dw,
               hipStream_t st) {
    const int key = tuning_key(n, ndens, sizeof(T) == 4);
    {
        std::lock_guard<std::mutex> lock(tuning().mu);
        auto it = tuning().shape.find(key);
        if (it != tuning().shape.end()) return it->second;
    }
    const size_t nn = (size_t)n * n;
    DevBuf<double> sJ, sK;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int best = kRows2;
    float best_ms = 0;
    bool ok = sJ.alloc(nn) == hipSuccess && sK.alloc(2 * nn) == hipSuccess &&
              hipEventCreate(&ev[0]) == hipSuccess && hipEventCreate(&ev[1]) == hipSuccess;
    for (int shape = 0; shape < kIncoreShapes; ++shape) {
        if (!ok) break;
        float t = 0;
        for (int rep = 0; rep < 3 && ok; ++rep) {          // first run warms the caches and the code
            ok = hipEventRecord(ev[0], st) == hipSuccess &&
                 launch_incore_shape(shape, ndens, G, n, np, D, Dw, sJ.p, sK.p, st) == hipSuccess &&
                 hipEventRecord(ev[1], st) == hipSuccess && hipEventSynchronize(ev[1]) == hipSuccess;
            float ms = 0;
            if (ok) ok = hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess;
            if (rep > 0) t = rep == 1 ? ms : std::min(t, ms);
        }
        if (ok && (best_ms == 0 || t < best_ms)) {
            best_ms = t;
            best = shape;
        }
    }
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipGetLastError();
    if (!ok) return kRows2;                                // tuning failed: the default, not recorded
    std::lock_guard<std::mutex> lock(tuning().mu);
    tuning().shape[key] = best;
    return best;
}

// Scratch of the allocation-free entry point (hf_build_jk_incore_device_noalloc): the packed weighted density
// Dw of one build, np = pair_pad_base(n) doubles, n <= 512.  One copy per GPU, for the whole process.  Host code
// never uses the array's name as an address: hf_incore_scratch_init takes the device address of each GPU's
// copy once, with that GPU current, and the entry point passes that address to the kernels.
constexpr int64_t kNoAllocMaxPairs = 131584;          // pair_pad_base(512): 1.0 MiB
__device__ __align__(256) double scratch[kNoAllocMaxPairs];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

int hf_incore_scratch_init(hf_ctx* ctx) {
    const int dev = ctx->opt.device;
    if (dev < 0 || dev >= kMaxDevices) return ctx->fail(HF_E_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev]) return HF_OK;
    HF_HIP_CHECK(ctx, hipSetDevice(dev));
    void* p = nullptr;
    HF_HIP_CHECK(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = (double*)p;
    return HF_OK;
}

void hf_release_incore(hf_ctx* ctx, size_t keep_bytes) {
    if (ctx->d_G.n * sizeof(double) + ctx->d_G32.n * sizeof(float) > keep_bytes) {
        ctx->d_G.release();
        ctx->d_G32.release();
    }
}

namespace {

// One in-core J/K build with the packed weighted density in Dw (np doubles).  no_alloc: the tile shape is the one
// already tuned (or the default), and no events are created or recorded, so nothing is allocated.
int incore_jk_run(hf_ctx* ctx, int ndens, const double* G, const float* g32, const double* d_D, double* d_J,
                  double* d_K, hipStream_t st, bool accumulate, double* Dw, bool no_alloc) {
    const int n = ctx->mol.nbf;
    const int64_t np = hfgpu::pair_pad_base(n);      // padded pair positions (rows and columns of the store)
    const size_t nn = (size_t)n * n;
    hipLaunchKernelGGL(incore_dw_kernel, dim3(n), dim3(256), 0, st, d_D, n, ndens, Dw);
    if (!accumulate) {
        hipLaunchKernelGGL(incore_fill_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, d_J, nn);
        hipLaunchKernelGGL(incore_fill_kernel, dim3((unsigned)((ndens * nn + 255) / 256)), dim3(256), 0, st, d_K,
                           ndens * nn);
    }
    // tile shape from the autotuner (first build of each kind in the process times the candidates)
    const int shape = no_alloc ? known_shape(n, ndens, g32 != nullptr)
                      : g32    ? tuned_shape(ctx, ndens, g32, n, np, d_D, Dw, st)
                               : tuned_shape(ctx, ndens, G, n, np, d_D, Dw, st);
    // statistics, and with profile = 1 events around the contraction (read at the next flush)
    const int f = g32 ? 1 : 0;
    const bool prof = ctx->opt.profile != 0 && !no_alloc;
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
        e = launch_incore_shape(shape, ndens, g32, n, np, d_D, Dw, d_J, d_K, st);
    else
        e = launch_incore_shape(shape, ndens, G, n, np, d_D, Dw, d_J, d_K, st);
    if (e != hipSuccess) return ctx->fail(HF_E_HIP, std::string("in-core J/K: ") + hipGetErrorString(e));
    if (prof) {
        HF_HIP_CHECK(ctx, hipEventRecord(ctx->incore_ev[1], st));
        ctx->incore_pending = f;
    }
    return HF_OK;
}

}  // namespace

int hf_incore_jk_impl(hf_ctx* ctx, int ndens, const double* G, const float* g32, const double* d_D, double* d_J,
                      double* d_K, hipStream_t st, bool accumulate) {
    if (!ctx->have_mol) return ctx->fail(HF_E_STATE, "no molecule set");
    if (ndens < 1 || ndens > 2) return ctx->fail(HF_E_ARG, "ndens must be 1 or 2");
    const int n = ctx->mol.nbf;
    if (n > kIncoreMaxN) return ctx->fail(HF_E_UNSUPPORTED, "in-core J/K takes at most 512 basis functions");
    HF_HIP_CHECK(ctx, ctx->d_Dw.alloc((size_t)hfgpu::pair_pad_base(n)));
    return incore_jk_run(ctx, ndens, G, g32, d_D, d_J, d_K, st, accumulate, ctx->d_Dw.p, false);
}

extern "C" int hf_build_jk_incore_device_noalloc(hf_ctx* ctx, int ndens, const double* d_G, const double* d_D,
                                                 double* d_J, double* d_K, void* stream) {
    if (!ctx) return HF_E_ARG;
    if (!ctx->have_mol) return ctx->fail(HF_E_STATE, "no molecule set");
    if (!d_G || !d_D || !d_J || !d_K) return ctx->fail(HF_E_ARG, "null device buffer");
    if (ndens < 1 || ndens > 2) return ctx->fail(HF_E_ARG, "ndens must be 1 or 2");
    const int n = ctx->mol.nbf;
    if (hfgpu::pair_pad_base(n) > kNoAllocMaxPairs)
        return ctx->fail(HF_E_ARG, "n = " + std::to_string(n) + " exceeds the allocation-free scratch (n <= 512)");
    const int dev = ctx->opt.device;
    double* Dw = (dev >= 0 && dev < kMaxDevices) ? g_scratch_ptr[dev] : nullptr;
    if (!Dw) return ctx->fail(HF_E_STATE, "allocation-free scratch not set up on this device");
    return incore_jk_run(ctx, ndens, d_G, nullptr, d_D, d_J, d_K, stream ? (hipStream_t)stream : ctx->stream, false,
                         Dw, true);
}
