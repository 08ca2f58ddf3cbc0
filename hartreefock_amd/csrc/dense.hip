// Dense FP64 kernels for the SCF step on small matrices (n up to a few hundred), where a
// library GEMM launches too few workgroups to use the chip:
//   * dgemm_small: C = alpha op(A) op(B) + beta C (column-major, as BLAS), one block of four
//     wavefronts per 16 x 16 tile of C, splitting k, on the FP64 matrix cores (v_mfma_f64_16x16x4_f64);
//   * the SP2 purification step fused into one such product (see sp2_step_kernel).
//
// Matrix-core lane layout of __builtin_amdgcn_mfma_f64_16x16x4f64 (A 16x4, B 4x16, D 16x16):
//   lane l holds A[l % 16][l / 16] and B[l / 16][l % 16]; element r of its 4-vector result is
//   D[l / 16 + 4 r][l % 16].
#include <mutex>

#include "gershgorin.h"
#include "hf_internal.h"
#include "wave_reduce.h"

namespace {

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

// One 16 x 16 tile of C per block of 4 wavefronts, which split the k range.
template <bool TA, bool TB>
__global__ void __launch_bounds__(256) dgemm16_kernel(int m, int n, int k, double alpha, const double* __restrict__ A,
                                                      int lda, const double* __restrict__ B, int ldb, double beta,
                                                      double* __restrict__ C, int ldc) {
    __shared__ v4d part[3][64];
    const int l = threadIdx.x & 63, lr = l & 15, lk = l >> 4, w = threadIdx.x >> 6;
    const int i0 = blockIdx.x * 16, j0 = blockIdx.y * 16;
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

// ---- SP2 purification, one kernel per step ----
// Step s forms X_s from the previous step's X_{s-1} and X_{s-1}^2 (or, for s = 0, from the
// orthogonalised Fock matrix and its spectral bounds) on the fly, writes X_s, and multiplies
// X_s X_s on the matrix cores; the wavefronts on diagonal tiles write partial traces of X_s and
// X_s^2 for the next step's decision.  Everything a step needs from the previous one is read
// after a kernel boundary, so one launch per step needs no host round trip.
//
// Step s (s >= 1) first decides, identically in every wavefront, for the previous matrix:
//   e_{s-1} = tr X_{s-1} - tr X_{s-1}^2,   square = |tr X^2 - nocc| <= |2 tr X - tr X^2 - nocc|,
//   stop_{s-1} = stop_{s-2} or (e_{s-1} < 1e-4 and e_{s-1} >= e_{s-3})
// (the stopping rule of Kruchinina, Rudberg and Rubensson, JCTC 12, 5788 (2016)).  If stop_{s-2}
// held, X_{s-1} was already final and the step only carries the stop flag forward.  Otherwise
// X_s = X_{s-1}^2 or
// 2 X_{s-1} - X_{s-1}^2 is written, and if stop_{s-1} holds it is the result (final = s).
//
// hist (device, doubles): [0] lo, [1] hi, [2] final step (-1 while running), then per step s at
// 8 + 2 s: e_s, stop_s.  part: per step s at 2 * nt * s: nt partial traces of X_s, then nt of X_s^2.
constexpr int kSp2Hist = 8;

__global__ void __launch_bounds__(1024) sp2_bounds_kernel(const double* __restrict__ F, int n, double* __restrict__ hist) {
    __shared__ double part[1024];
    double lo, hi;
    hfgpu::gershgorin_block(F, n, part, lo, hi);
    if (threadIdx.x == 0) {
        if (hi - lo < 1e-12) hi = lo + 1.0;
        hist[0] = lo;
        hist[1] = hi;
        hist[2] = -1.0;
    }
}

struct Sp2Args {
    int n, nt, s;
    double nocc;
    const double* F;         // orthogonalised Fock matrix (step 0)
    const double* Xp;        // X_{s-1}
    const double* X2p;       // X_{s-1}^2
    double* X;               // X_s (out)
    double* X2;              // X_s^2 (out)
    double* hist;
    double* part;
};

// Operand loads of one wavefront's k strip, issued before the step's decision is read (the decision
// needs the previous step's partial traces, one more round trip): up to kSp2Trips x 8 k values.
constexpr int kSp2Trips = ((kSmallDense + 15) / 16 * 4 + 31) / 32;

__global__ void __launch_bounds__(256) sp2_step_kernel(Sp2Args g) {
    __shared__ v4d part4[3][64];
    const int n = g.n, nt = g.nt, s = g.s;
    const int l = threadIdx.x & 63, lr = l & 15, lk = l >> 4, w = threadIdx.x >> 6;
    const int i0 = blockIdx.x * 16, j0 = blockIdx.y * 16;
    const bool lead = blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0;
    int kb, ke;
    k_strip(n, w, kb, ke);
    // raw operands: element (i, kk) of A and (j, kk) of B (X_s symmetric), from X_{s-1} and
    // X_{s-1}^2, or from F for s = 0
    double ra[kSp2Trips][8], ra2[kSp2Trips][8], rb[kSp2Trips][8], rb2[kSp2Trips][8];
    const double* P1 = s == 0 ? g.F : g.Xp;
    const double* P2 = s == 0 ? g.F : g.X2p;
#pragma unroll
    for (int q = 0; q < kSp2Trips; ++q)
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int kk = kb + 32 * q + 4 * u + lk, ia = i0 + lr, jb = j0 + lr;
            const bool okk = kk < ke;
            const size_t oa = ia + (size_t)kk * n, ob = jb + (size_t)kk * n;
            ra[q][u] = okk && ia < n ? P1[oa] : 0.0;
            ra2[q][u] = okk && ia < n && s > 0 ? P2[oa] : 0.0;
            rb[q][u] = okk && jb < n ? P1[ob] : 0.0;
            rb2[q][u] = okk && jb < n && s > 0 ? P2[ob] : 0.0;
        }
    bool square = false;
    if (s > 0) {
        const double* pp = g.part + (size_t)2 * nt * (s - 1);
        double t1 = 0, t2 = 0;
        for (int i = 0; i < nt; ++i) {
            t1 += pp[i];
            t2 += pp[nt + i];
        }
        const double e = t1 - t2;
        square = fabs(t2 - g.nocc) <= fabs(2.0 * t1 - t2 - g.nocc);
        const bool stop_pp = s >= 2 && g.hist[kSp2Hist + 2 * (s - 2) + 1] != 0;
        if (stop_pp) {                                         // X_{s-1} already final
            if (lead) g.hist[kSp2Hist + 2 * (s - 1) + 1] = 1.0;
            return;
        }
        const double e3 = s >= 3 ? g.hist[kSp2Hist + 2 * (s - 3)] : 1e300;
        const bool stop_p = e < 1e-4 && e >= e3;
        if (lead) {
            g.hist[kSp2Hist + 2 * (s - 1)] = e;
            g.hist[kSp2Hist + 2 * (s - 1) + 1] = stop_p ? 1.0 : 0.0;
            if (stop_p) g.hist[2] = s;
        }
        if (stop_p) {
            // X_s is the result: write this tile of it and stop
            if (w == 0)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = i0 + lk + 4 * r, j = j0 + lr;
                    if (i < n && j < n) {
                        const size_t o = i + (size_t)j * n;
                        g.X[o] = square ? g.X2p[o] : 2.0 * g.Xp[o] - g.X2p[o];
                    }
                }
            return;
        }
    }
    const double lo = g.hist[0], hi = g.hist[1], inv = 1.0 / (hi - lo);
    auto combine = [&](double x, double x2, int i, int kk) -> double {   // element of X_s from the raw values
        if (s == 0) return ((i == kk ? hi : 0.0) - x) * inv;
        return square ? x2 : 2.0 * x - x2;
    };
    v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < kSp2Trips; ++q)
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int kk = kb + 32 * q + 4 * u + lk, ia = i0 + lr, jb = j0 + lr;
            const bool oka = kk < ke && ia < n, okb = kk < ke && jb < n;
            const double a = oka ? combine(ra[q][u], ra2[q][u], ia, kk) : 0.0;
            const double b = okb ? combine(rb[q][u], rb2[q][u], jb, kk) : 0.0;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
        }
    acc = reduce_waves(acc, part4);
    if (w != 0) return;
    auto xval = [&](int i, int j) -> double {                   // element (i, j) of X_s, zero outside
        if (i >= n || j >= n) return 0.0;
        const size_t o = i + (size_t)j * n;
        if (s == 0) return ((i == j ? hi : 0.0) - g.F[o]) * inv;
        return square ? g.X2p[o] : 2.0 * g.Xp[o] - g.X2p[o];
    };
    double d1 = 0, d2 = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = i0 + lk + 4 * r, j = j0 + lr;
        if (i < n && j < n) {
            const size_t o = i + (size_t)j * n;
            const double x = xval(i, j);
            g.X[o] = x;
            g.X2[o] = acc[r];
            if (i == j) {
                d1 += x;
                d2 += acc[r];
            }
        }
    }
    if (i0 == j0) {   // diagonal tile: partial traces, reduced over the wavefront in a fixed order
        double V[2] = {d1, d2};
        hfgpu::Halving<2, 32>::run(V, l);     // lanes 0..31 now hold the d1 total, lanes 32..63 the d2 total
        double* pp = g.part + (size_t)2 * nt * s;
        if (l == 0) pp[blockIdx.x] = V[0];
        if (l == 32) pp[nt + blockIdx.x] = V[0];
    }
}

// SP2 for n <= kSp2LocalMax in one workgroup: X and X^2 stay in LDS for every step, so the whole
// purification is one launch with no host round trip (small molecules, served many at a time).
// Same steps and stopping rule as sp2_step_kernel; 16 wavefronts share the 16 x 16 tiles of each
// product on the matrix cores.  out[0] = final step (or -1 if the rule was not met), then X.
constexpr int kSp2LocalMax = 96;

__global__ void __launch_bounds__(1024) sp2_local_kernel(const double* __restrict__ Fp, int n, double nocc,
                                                          int max_steps, double* __restrict__ out,
                                                          double* __restrict__ Xout) {
    extern __shared__ double lds[];
    double* X = lds;              // n x n, column-major (symmetric)
    double* X2 = lds + n * n;
    __shared__ double red[2][32];
    __shared__ double e_hist[3];
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int l = tid & 63, lr = l & 15, lk = l >> 4, w = tid >> 6, nw = nthr >> 6;
    // Gershgorin bounds (16 row groups per column, so each thread reads ~n/16 elements)
    double lo, hi;
    hfgpu::gershgorin_block(Fp, n, lds, lo, hi);      // (X and X^2 are not set up yet)
    if (hi - lo < 1e-12) hi = lo + 1.0;
    const double inv = 1.0 / (hi - lo);
    for (int o = tid; o < n * n; o += nthr) X[o] = ((o / n == o % n ? hi : 0.0) - Fp[o]) * inv;
    if (tid == 0) e_hist[0] = e_hist[1] = e_hist[2] = 1e300;   // e_{s-1}, e_{s-2}
    __syncthreads();
    const int nt = (n + 15) / 16;
    int final_step = -1;
    for (int s = 0; s <= max_steps; ++s) {
        // X2 = X X, one 16 x 16 tile per wavefront in turn
        for (int tile = w; tile < nt * nt; tile += nw) {
            const int i0 = (tile % nt) * 16, j0 = (tile / nt) * 16;
            v4d acc = {0.0, 0.0, 0.0, 0.0};
            for (int k0 = 0; k0 < n; k0 += 4) {
                const int kk = k0 + lk, ia = i0 + lr, jb = j0 + lr;
                const double a = kk < n && ia < n ? X[ia + kk * n] : 0.0;
                const double b = kk < n && jb < n ? X[jb + kk * n] : 0.0;
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + lk + 4 * r, j = j0 + lr;
                if (i < n && j < n) X2[i + j * n] = acc[r];
            }
        }
        __syncthreads();
        // traces, in a fixed order
        double t1 = 0, t2 = 0;
        for (int i = tid; i < n; i += nthr) {
            t1 += X[i + i * n];
            t2 += X2[i + i * n];
        }
        {
            double V[2] = {t1, t2};
            hfgpu::Halving<2, 32>::run(V, l);  // lanes 0..31: the t1 total, lanes 32..63: the t2 total
            if (l == 0) red[0][w] = V[0];
            if (l == 32) red[1][w] = V[0];
        }
        __syncthreads();
        t1 = 0;
        t2 = 0;
        for (int q = 0; q < nw; ++q) {
            t1 += red[0][q];
            t2 += red[1][q];
        }
        const double e = t1 - t2;
        const bool square = fabs(t2 - nocc) <= fabs(2.0 * t1 - t2 - nocc);
        const bool stop = e < 1e-4 && e >= e_hist[1];
        for (int o = tid; o < n * n; o += nthr) X[o] = square ? X2[o] : 2.0 * X[o] - X2[o];
        __syncthreads();
        if (tid == 0) {
            e_hist[1] = e_hist[0];
            e_hist[0] = e;
        }
        __syncthreads();
        if (stop) {
            final_step = s + 1;
            break;
        }
    }
    for (int o = tid; o < n * n; o += nthr) Xout[o] = X[o];
    if (tid == 0) out[0] = final_step;
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

namespace {

// The dynamic-LDS limit of sp2_local_kernel is an attribute of the kernel for the whole process, not of
// one launch.  It is set once, to the size of the largest matrix the kernel takes (n = kSp2LocalMax), before
// any launch: setting it on every call to that call's size, from the concurrent contexts of a server's
// slots, changed it under the other threads' launches of the same kernel.
bool sp2_local_ready() {
    static std::once_flag once;
    static bool ok = false;
    std::call_once(once, [] {
        const size_t lds = 2 * (size_t)kSp2LocalMax * kSp2LocalMax * sizeof(double);
        ok = hipFuncSetAttribute((const void*)sp2_local_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)lds) == hipSuccess;
        (void)hipGetLastError();
    });
    return ok;
}

__global__ void set_flag_kernel(double* __restrict__ p, double v) { *p = v; }

}  // namespace

int hf_sp2_fused(hf_ctx* ctx, hipStream_t st, int n, int nocc, const double* Fp, double* X0, double* X1, double* W0,
                 double* W1, double* hist, double* part, int max_steps, int hint, double*& result, int& steps) {
    const int nt = (n + 15) / 16;
    if (n <= kSp2LocalMax && sp2_local_ready()) {
        const size_t lds = std::max<size_t>(2 * (size_t)n * n, 1024) * sizeof(double);   // >= the bounds' workspace
        // the stop flag starts at "not met", so nothing a launch failed to write is read as a result
        hipLaunchKernelGGL(set_flag_kernel, dim3(1), dim3(1), 0, st, hist + 2, -1.0);
        hipLaunchKernelGGL(sp2_local_kernel, dim3(1), dim3(1024), lds, st, Fp, n, (double)nocc, max_steps, hist + 2, X0);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return ctx->fail(HF_E_HIP, std::string("single-workgroup SP2 launch: ") + hipGetErrorString(e));
        double fin = -1;
        if (int rc = hf_readback(ctx, st, &fin, hist + 2, 1)) return rc;
        result = fin >= 0 ? X0 : nullptr;
        steps = (int)fin;
        return HF_OK;
    }
    // the stop flag is read back from step kFirstCheck on (or from the step count of the previous
    // call, hint, when given) and then every kEvery steps
    const int kFirstCheck = hint > 0 ? std::max(8, hint) : 16, kEvery = hint > 0 ? 2 : 4;
    hipLaunchKernelGGL(sp2_bounds_kernel, dim3(1), dim3(1024), 0, st, Fp, n, hist);
    Sp2Args g{};
    g.n = n;
    g.nt = nt;
    g.nocc = nocc;
    g.F = Fp;
    g.hist = hist;
    g.part = part;
    double* Xb[2] = {X0, X1};
    double* Wb[2] = {W0, W1};
    result = nullptr;
    steps = 0;
    for (int s = 0; s <= max_steps; ++s) {
        g.s = s;
        g.Xp = Xb[(s + 1) & 1];
        g.X2p = Wb[(s + 1) & 1];
        g.X = Xb[s & 1];
        g.X2 = Wb[s & 1];
        hipLaunchKernelGGL(sp2_step_kernel, dim3(nt, nt), dim3(256), 0, st, g);
        if ((s >= kFirstCheck && (s - kFirstCheck) % kEvery == 0) || s == max_steps) {
            double fin = -1;
            if (int rc = hf_readback(ctx, st, &fin, hist + 2, 1)) return rc;
            if (fin >= 0) {
                steps = (int)fin;
                result = Xb[steps & 1];
                break;
            }
        }
    }
    HF_HIP_CHECK(ctx, hipGetLastError());
    return HF_OK;
}
