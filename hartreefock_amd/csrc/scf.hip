// Device-resident SCF (RHF / UHF) around the direct J/K build.
//
// Control flow and numerics follow the reference library:
//   HartreeFockAlgorithm::Init / Calculate / First-, SelfConsistent-, NormalIterations
//     (HartreeFock/HartreeFockAlgorithm.cpp:21-192),
//   RestrictedHartreeFock::Step / DIISStep / InitFockMatrix / CalculateEnergy
//     (HartreeFock/RestrictedHartreeFock.cpp:35-249),
//   UnrestrictedHartreeFock::Step / DIISStep / InitFockMatrices / CalculateEnergy
//     (HartreeFock/UnrestrictedHartreeFock.cpp:62-293),
//   DIIS (HartreeFock/DIIS.h:9-96): 6 kept, extrapolation from 5, B c = [0..0 1].
// All matrices live in HBM; the dense steps use rocBLAS dgemm (FP64).  The reference
// diagonalises the orthogonalised Fock matrix every step; here the new density is by
// default the same aufbau projector obtained by SP2 purification (dgemm only), with
// rocSOLVER dsyevd as the fallback and for the final orbital energies.  Only a few
// scalars and the DIIS inner products come back to the host each iteration.  Matrices
// are symmetric except the MO coefficients C, which are stored column-major (column
// j = orbital j) as rocSOLVER returns them.
#include <rocsolver/rocsolver.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <deque>
#include <limits>

#include "gershgorin.h"
#include "scf_run.h"
#include "wave_reduce.h"

namespace {

__global__ void add_kernel(const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ c,
                           size_t nn) {
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nn; idx += (size_t)gridDim.x * blockDim.x)
        c[idx] = a[idx] + b[idx];
}

__global__ void to_f32_kernel(const double* __restrict__ a, float* __restrict__ b, size_t n) {
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x)
        b[idx] = (float)a[idx];
}

// F = g S_ij (h_ii + h_jj)/2 (g > 0) or h; Fa = F plus `asym` on (0,1) and (1,0).
__global__ void guess_kernel(const double* __restrict__ S, const double* __restrict__ h, double g, int n, double asym,
                             double* __restrict__ Fa, double* __restrict__ Fb) {
    const size_t nn = (size_t)n * n;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nn; idx += (size_t)gridDim.x * blockDim.x) {
        const size_t i = idx / n, j = idx % n;
        const double f = g > 0 ? g * S[idx] * (h[i * n + i] + h[j * n + j]) * 0.5 : h[idx];
        Fb[idx] = f;
        Fa[idx] = (idx == 1 || idx == (size_t)n) ? f + asym : f;
    }
}

__global__ void sym_kernel(double* __restrict__ A, int n) {
    const size_t nn = (size_t)n * n;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nn; idx += (size_t)gridDim.x * blockDim.x) {
        const size_t i = idx % n, j = idx / n;
        if (i < j) {
            const double v = 0.5 * (A[idx] + A[i * n + j]);
            A[idx] = v;
            A[i * n + j] = v;
        }
    }
}

// A = c B, and I (identity) into E
__global__ void scale_ident_kernel(const double* __restrict__ B, double c, double* __restrict__ A,
                                   double* __restrict__ E, int n) {
    const size_t nn = (size_t)n * n;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nn; idx += (size_t)gridDim.x * blockDim.x) {
        A[idx] = c * B[idx];
        E[idx] = idx / n == idx % n ? 1.0 : 0.0;
    }
}

// T <- (3 I - T) / 2
__global__ void ns_step_kernel(double* __restrict__ T, int n) {
    const size_t nn = (size_t)n * n;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nn; idx += (size_t)gridDim.x * blockDim.x)
        T[idx] = 0.5 * ((idx / n == idx % n ? 3.0 : 0.0) - T[idx]);
}

// W[:, j] = U[:, j] / sqrt(s_j)   (column-major)
__global__ void scale_cols_kernel(const double* __restrict__ U, const double* __restrict__ s, double* __restrict__ W,
                                  int n) {
    const size_t nn = (size_t)n * n;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nn; idx += (size_t)gridDim.x * blockDim.x) {
        const size_t j = idx / n;
        W[idx] = U[idx] / sqrt(s[j]);
    }
}

// out = sum_k c[k] * slots[idx[k]]
struct LinComb {
    const double* src[8];
    double c[8];
    int m;
};
__global__ void lincomb_kernel(LinComb lc, double* __restrict__ out, size_t nn) {
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nn; idx += (size_t)gridDim.x * blockDim.x) {
        double v = 0;
        for (int k = 0; k < lc.m; ++k) v += lc.c[k] * lc.src[k][idx];
        out[idx] = v;
    }
}

// Deterministic reductions: out[r] = sum_i x_r[i] * y_r[i]  or  sum_i (x_r[i]-y_r[i])^2, r < m <= 8.
// kDotBlocks blocks each reduce one contiguous chunk of every job into part[r * kDotBlocks + b]; one
// wavefront then sums the partials of each job in a fixed order.
struct DotJobs {
    const double* x[8];
    const double* y[8];
    int diff[8];
    int m;
};
constexpr int kDotBlocks = 64;

__global__ void __launch_bounds__(256) dots_partial_kernel(DotJobs jobs, size_t nn, double* __restrict__ part) {
    __shared__ double sh[256];
    const size_t chunk = (nn + gridDim.x - 1) / gridDim.x;
    const size_t lo = (size_t)blockIdx.x * chunk, hi = lo + chunk < nn ? lo + chunk : nn;
    for (int r = 0; r < jobs.m; ++r) {
        double acc = 0;
        const double* __restrict__ x = jobs.x[r];
        const double* __restrict__ y = jobs.y[r];
        if (jobs.diff[r]) {
            for (size_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
                const double d = x[i] - y[i];
                acc += d * d;
            }
        } else {
            for (size_t i = lo + threadIdx.x; i < hi; i += blockDim.x) acc += x[i] * y[i];
        }
        sh[threadIdx.x] = acc;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[r * gridDim.x + blockIdx.x] = sh[0];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(64) dots_final_kernel(const double* __restrict__ part, int m, double* __restrict__ out) {
    for (int r = 0; r < m; ++r) {
        double v[1] = {part[r * kDotBlocks + threadIdx.x]};
        hfgpu::Halving<1, 32>::run(v, (int)threadIdx.x);
        if (threadIdx.x == 0) out[r] = v[0];
    }
}

// part: kDotBlocks * 8 doubles of workspace
void launch_dots(hipStream_t s, const DotJobs& jobs, size_t nn, double* part, double* out) {
    hipLaunchKernelGGL(dots_partial_kernel, dim3(kDotBlocks), dim3(256), 0, s, jobs, nn, part);
    hipLaunchKernelGGL(dots_final_kernel, dim3(1), dim3(64), 0, s, part, jobs.m, out);
}

int grid_for(size_t n) {
    size_t b = (n + 255) / 256;
    return (int)std::min<size_t>(std::max<size_t>(b, 1), 4096);
}

// ---- Density matrix by second-order spectral projection (SP2; Niklasson, PRB 66, 155115 (2002)) ----
// For the orthogonalised Fock matrix F' the aufbau density is the spectral projector onto its
// `nocc` lowest eigenvectors.  Starting from X0 = (hi I - F')/(hi - lo), with [lo, hi] enclosing the
// spectrum (Gershgorin discs), every step replaces X by X^2 or 2X - X^2, whichever keeps trace(X)
// closer to nocc; the eigenvalues are driven to 0 or 1 with the eigenvectors unchanged.  Each step
// is one dgemm plus two small kernels, and no eigendecomposition is needed.  Stopping rule
// (Kruchinina, Rudberg, Rubensson, JCTC 12, 5788 (2016)): once the idempotency error
// e_k = trace(X - X^2) is small, stop as soon as it no longer decreases (e_k >= e_{k-2}).
// State (device, 8 doubles): lo, hi, done, e_{k-1}, e_{k-2}, steps, use_square, apply.
enum { SP2_LO, SP2_HI, SP2_DONE, SP2_E1, SP2_E2, SP2_STEPS, SP2_SQUARE, SP2_APPLY };

__device__ double block_sum_256(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(1024) sp2_bounds_kernel(const double* __restrict__ F, int n, double* __restrict__ st) {
    __shared__ double part[1024];
    double lo, hi;
    hfgpu::gershgorin_block(F, n, part, lo, hi);
    if (threadIdx.x == 0) {
        if (hi - lo < 1e-12) hi = lo + 1.0;
        st[SP2_LO] = lo;
        st[SP2_HI] = hi;
        st[SP2_DONE] = 0;
        st[SP2_E1] = 1e300;
        st[SP2_E2] = 1e300;
        st[SP2_STEPS] = 0;
        st[SP2_SQUARE] = 0;
        st[SP2_APPLY] = 0;
    }
}

__global__ void sp2_init_kernel(const double* __restrict__ F, double* __restrict__ X, int n,
                                const double* __restrict__ st) {
    const size_t nn = (size_t)n * n;
    const double lo = st[SP2_LO], hi = st[SP2_HI], inv = 1.0 / (hi - lo);
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nn; idx += (size_t)gridDim.x * blockDim.x) {
        const size_t i = idx / n, j = idx % n;
        X[idx] = ((i == j ? hi : 0.0) - F[idx]) * inv;
    }
}

__global__ void __launch_bounds__(256) sp2_decide_kernel(const double* __restrict__ X, const double* __restrict__ X2,
                                                         int n, double nocc, double* __restrict__ st) {
    __shared__ double sh[256];
    if (st[SP2_DONE] != 0) {
        if (threadIdx.x == 0) st[SP2_APPLY] = 0;
        return;
    }
    double t1 = 0, t2 = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        t1 += X[(size_t)i * n + i];
        t2 += X2[(size_t)i * n + i];
    }
    t1 = block_sum_256(t1, sh);
    t2 = block_sum_256(t2, sh);
    if (threadIdx.x == 0) {
        const double e = t1 - t2;
        st[SP2_SQUARE] = fabs(t2 - nocc) <= fabs(2.0 * t1 - t2 - nocc) ? 1.0 : 0.0;
        st[SP2_APPLY] = 1.0;
        st[SP2_STEPS] += 1.0;
        if (e < 1e-4 && e >= st[SP2_E2]) st[SP2_DONE] = 1.0;
        st[SP2_E2] = st[SP2_E1];
        st[SP2_E1] = e;
    }
}

__global__ void sp2_update_kernel(double* __restrict__ X, const double* __restrict__ X2, size_t nn,
                                  const double* __restrict__ st) {
    if (st[SP2_APPLY] == 0) return;
    const bool sq = st[SP2_SQUARE] != 0;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nn; idx += (size_t)gridDim.x * blockDim.x)
        X[idx] = sq ? X2[idx] : 2.0 * X[idx] - X2[idx];
}

#define BL(call)                                                                                        \
    do {                                                                                                \
        rocblas_status s_ = (call);                                                                     \
        if (s_ != rocblas_status_success)                                                               \
            return ctx->fail(HF_E_HIP, std::string(#call) + ": rocblas status " + std::to_string((int)s_)); \
    } while (0)

// Products of Init and finish on the context's stream.
int gemm(hf_ctx* ctx, bool ta, bool tb, int m, int n, int k, double alpha, const double* A, const double* B,
         double beta, double* C) {
    return scf_gemm(ctx, ctx->stream, ta, tb, m, n, k, alpha, A, B, beta, C);
}

// S^{-1/2} into st.X.  The reference diagonalises S (HF/HartreeFockAlgorithm.cpp:48-59); the
// symmetric inverse square root is unique, and here it comes from the coupled Newton-Schulz
// iteration (Higham, Numer. Algorithms 15, 227 (1997)) on Y0 = c S, c = 1/||S||_F (spectrum in
// (0, 1]):  T = (3I - Z Y)/2,  Y <- Y T,  Z <- T Z;  Z -> (cS)^{-1/2}, so S^{-1/2} = sqrt(c) Z.
// Three products per step (small matrices: matrix-core kernel); converged when ||I - Z Y||_F is
// below 1e-14 sqrt(n) or, once below 1e-8, stops halving (the rounding floor of an ill-conditioned
// S); if neither happens within 80 steps (S numerically singular), dsyevd.
int inverse_sqrt(hf_ctx* ctx, ScfState& st) {
    hipStream_t s = ctx->stream;
    const int n = st.n;
    const size_t nn = st.nn;
    DotJobs jobs{};
    jobs.m = 1;
    jobs.x[0] = st.S.p;
    jobs.y[0] = st.S.p;
    launch_dots(s, jobs, nn, st.redpart.p, st.red.p);
    double fro2 = 0;
    if (int rc = hf_readback(ctx, s, &fro2, st.red.p, 1)) return rc;
    const double c = 1.0 / std::sqrt(fro2);
    double *Y = st.T1.p, *Z = st.X.p, *T = st.T2.p, *W = st.T3.p, *E = st.err.p;
    hipLaunchKernelGGL(scale_ident_kernel, dim3(grid_for(nn)), dim3(256), 0, s, st.S.p, c, Y, Z, n);
    HF_HIP_CHECK(ctx, hipMemcpyAsync(E, Z, nn * sizeof(double), hipMemcpyDeviceToDevice, s));
    bool ok = false;
    double prev = 1e300;
    for (int it = 1; it <= 80 && !ok; ++it) {
        if (int rc = gemm(ctx, false, false, n, n, n, 1.0, Z, Y, 0.0, T)) return rc;
        if (it % 4 == 0 || it > 24) {
            jobs.x[0] = T;
            jobs.y[0] = E;
            jobs.diff[0] = 1;
            launch_dots(s, jobs, nn, st.redpart.p, st.red.p);
            double r2 = 0;
            if (int rc = hf_readback(ctx, s, &r2, st.red.p, 1)) return rc;
            const double r = std::sqrt(r2);
            if (r < 1e-14 * std::sqrt((double)n) || (r < 1e-8 && r >= 0.5 * prev)) ok = true;
            prev = r;
        }
        hipLaunchKernelGGL(ns_step_kernel, dim3(grid_for(nn)), dim3(256), 0, s, T, n);
        if (int rc = gemm(ctx, false, false, n, n, n, 1.0, Y, T, 0.0, W)) return rc;
        std::swap(Y, W);
        if (int rc = gemm(ctx, false, false, n, n, n, 1.0, T, Z, 0.0, W)) return rc;
        std::swap(Z, W);
    }
    if (ok) {
        hipLaunchKernelGGL(scale_ident_kernel, dim3(grid_for(nn)), dim3(256), 0, s, Z, std::sqrt(c), st.C.p, T, n);
        HF_HIP_CHECK(ctx, hipMemcpyAsync(st.X.p, st.C.p, nn * sizeof(double), hipMemcpyDeviceToDevice, s));
    } else {
        // S^{-1/2} = U s^{-1/2} U^T
        HF_HIP_CHECK(ctx, hipMemcpyAsync(st.T1.p, st.S.p, nn * sizeof(double), hipMemcpyDeviceToDevice, s));
        if (int rc = scf_syevd(ctx, st, ctx->stream, st.T1.p, st.eps.p)) return rc;
        hipLaunchKernelGGL(scale_cols_kernel, dim3(grid_for(nn)), dim3(256), 0, s, st.T1.p, st.eps.p, st.T2.p, n);
        if (int rc = gemm(ctx, false, true, n, n, n, 1.0, st.T2.p, st.T1.p, 0.0, st.X.p)) return rc;
    }
    hipLaunchKernelGGL(sym_kernel, dim3(grid_for(nn)), dim3(256), 0, s, st.X.p, n);
    HF_HIP_CHECK(ctx, hipGetLastError());
    return HF_OK;
}

// Aufbau projector of the orthogonalised Fock matrix Fp (not modified) onto its nocc lowest
// eigenvectors, by SP2 steps, into P; W is workspace.  ok = false when the stopping rule was not
// met within kMaxSteps (a vanishing HOMO-LUMO gap leaves fractional eigenvalues).
}  // namespace

int scf_sp2_projector(hf_ctx* ctx, ScfState& st, hipStream_t s, const double* Fp, int nocc, double* P, double* W,
                      bool& ok) {
    const int n = st.n;
    constexpr int kMaxSteps = 100, kFirstCheck = 16, kEvery = 4;
    hipLaunchKernelGGL(sp2_bounds_kernel, dim3(1), dim3(1024), 0, s, Fp, n, st.sp2.p);
    hipLaunchKernelGGL(sp2_init_kernel, dim3(grid_for(st.nn)), dim3(256), 0, s, Fp, P, n, st.sp2.p);
    ok = false;
    for (int k = 1; k <= kMaxSteps; ++k) {
        if (int rc = scf_gemm(ctx, s, false, false, n, n, n, 1.0, P, P, 0.0, W)) return rc;
        hipLaunchKernelGGL(sp2_decide_kernel, dim3(1), dim3(256), 0, s, P, W, n, (double)nocc, st.sp2.p);
        hipLaunchKernelGGL(sp2_update_kernel, dim3(grid_for(st.nn)), dim3(256), 0, s, P, W, st.nn, st.sp2.p);
        if ((k >= kFirstCheck && (k - kFirstCheck) % kEvery == 0) || k == kMaxSteps) {
            double done = 0;
            if (int rc = hf_readback(ctx, s, &done, st.sp2.p + SP2_DONE, 1)) return rc;
            if (done != 0) {
                ok = true;
                break;
            }
        }
    }
    HF_HIP_CHECK(ctx, hipGetLastError());
    return HF_OK;
}

// Symmetric eigenproblem, in place: A -> eigenvectors (column-major), w ascending, on stream st.
// HFGPU_EIGEN selects the rocSOLVER routine: syevd (divide and conquer, default), syevj (Jacobi),
// syevdj, syev (QR iteration).  All FP64.
int scf_syevd(hf_ctx* ctx, ScfState& st, hipStream_t stream, double* A, double* w) {
    static const char* env = std::getenv("HFGPU_EIGEN");
    static const int alg = !env ? 0 : (!std::strcmp(env, "syevj") ? 1 : !std::strcmp(env, "syevdj") ? 2
                                                                     : !std::strcmp(env, "syev") ? 3 : 0);
    if (stream != ctx->stream) rocblas_set_stream(ctx->blas, stream);
    rocblas_status s;
    if (alg == 1)
        s = rocsolver_dsyevj(ctx->blas, rocblas_esort_ascending, rocblas_evect_original, rocblas_fill_upper, st.n, A,
                             st.n, 0.0, st.red.p + 8, 100, st.info.p + 1, w, st.info.p);
    else if (alg == 2)
        s = rocsolver_dsyevdj(ctx->blas, rocblas_evect_original, rocblas_fill_upper, st.n, A, st.n, w, st.info.p);
    else if (alg == 3)
        s = rocsolver_dsyev(ctx->blas, rocblas_evect_original, rocblas_fill_upper, st.n, A, st.n, w, st.work.p,
                            st.info.p);
    else
        s = rocsolver_dsyevd(ctx->blas, rocblas_evect_original, rocblas_fill_upper, st.n, A, st.n, w, st.work.p,
                             st.info.p);
    if (stream != ctx->stream) rocblas_set_stream(ctx->blas, ctx->stream);
    if (s != rocblas_status_success) return ctx->fail(HF_E_HIP, "rocsolver eigensolver: status " + std::to_string((int)s));
    return HF_OK;
}

ScfWorkspace::~ScfWorkspace() {
    for (hipEvent_t e : jk_ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : eri_ev)
        if (e) (void)hipEventDestroy(e);
    if (ready_ev) (void)hipEventDestroy(ready_ev);
}

void hf_scf_workspace_free(hf_ctx* ctx) {
    delete ctx->scf_ws;
    ctx->scf_ws = nullptr;
}

double hf_ctx_held_bytes(const hf_ctx* ctx) {
    double b = (double)ctx->d_G.n * 8.0 + (double)ctx->d_G32.n * 4.0 + (double)ctx->d_Dw.n * 8.0 +
               8.0 * (double)(ctx->d_Jp.n + ctx->d_Kp.n + ctx->d_D.n + ctx->d_J.n + ctx->d_K.n + ctx->d_fx.n);
    if (const ScfWorkspace* w = ctx->scf_ws) {
        const ScfState& st = w->st;
        for (const DevBuf<double>* m : {&st.S, &st.h, &st.X, &st.F0, &st.Fa, &st.Fb, &st.J, &st.K, &st.D, &st.Dn,
                                        &st.Dprev, &st.dD, &st.C, &st.eps, &st.T1, &st.T2, &st.T3, &st.err, &st.work,
                                        &st.red, &st.redpart, &st.sp2, &st.sp2W, &w->diisF[0],
                                        &w->diisF[1], &w->diisE[0], &w->diisE[1]})
            b += 8.0 * (double)m->n;
    }
    return b;
}


int ScfRun::init() {
    t_start = std::chrono::steady_clock::now();
    s = ctx->stream;
    uhf = p->unrestricted != 0;
    if (!ctx->scf_ws) ctx->scf_ws = new (std::nothrow) ScfWorkspace();
    if (!ctx->scf_ws) return ctx->fail(HF_E_OOM, "SCF workspace");
    ScfWorkspace& ws = *ctx->scf_ws;
    ScfState& st = ws.st;
    n = ctx->mol.nbf;
    st.n = n;
    st.nn = (size_t)n * n;
    nn = st.nn;
    // the workspace matrices are allocated for at least 256 functions, so a context that solves one molecule after
    // another (a server's slot) does not regrow them for every larger molecule
    const size_t nc = std::max<size_t>(nn, (size_t)kWorkspaceMinN * kWorkspaceMinN);
    // ---- Init (HartreeFock/HartreeFockAlgorithm.cpp:21-101) ----
    Enuc = hfhost::nuclear_repulsion(ctx->mol);
    nsp = uhf ? 2 : 1;
    HF_HIP_CHECK(ctx, st.S.alloc(nc));
    HF_HIP_CHECK(ctx, st.h.alloc(nc));
    HF_HIP_CHECK(ctx, st.X.alloc(nc));
    HF_HIP_CHECK(ctx, st.F0.alloc(2 * nc));
    HF_HIP_CHECK(ctx, st.Fa.alloc(2 * nc));
    HF_HIP_CHECK(ctx, st.J.alloc(nc));
    HF_HIP_CHECK(ctx, st.K.alloc(2 * nc));
    HF_HIP_CHECK(ctx, st.D.alloc(2 * nc));
    HF_HIP_CHECK(ctx, st.Dn.alloc(2 * nc));
    HF_HIP_CHECK(ctx, st.Dprev.alloc(2 * nc));
    HF_HIP_CHECK(ctx, st.dD.alloc(2 * nc));
    HF_HIP_CHECK(ctx, st.C.alloc(2 * nc));
    HF_HIP_CHECK(ctx, st.eps.alloc(2 * n));
    HF_HIP_CHECK(ctx, st.T1.alloc(nc));
    HF_HIP_CHECK(ctx, st.T2.alloc(nc));
    HF_HIP_CHECK(ctx, st.T3.alloc(nc));
    HF_HIP_CHECK(ctx, st.sp2.alloc(8 + 2 * (kSp2MaxSteps + 1)));
    if (n <= kSmallDense) HF_HIP_CHECK(ctx, st.sp2W.alloc(2 * nc));
    HF_HIP_CHECK(ctx, st.err.alloc(nc));
    HF_HIP_CHECK(ctx, st.work.alloc(n));
    HF_HIP_CHECK(ctx, st.red.alloc(16));
    HF_HIP_CHECK(ctx, st.redpart.alloc(8 * kDotBlocks));
    HF_HIP_CHECK(ctx, st.info.alloc(2));
    // S, T, V on the device; h = T + V
    // in-core ERIs (the reference's store-once scheme) when the pair matrix fits the budget; the
    // mixed-precision tier contracts an FP32 copy until its final FP64 phase.  The store is evaluated
    // on its own stream, overlapping the one-electron integrals, S^-1/2 and iteration 0 (which needs
    // no J/K); the first J/K build waits for it.
    mixed_tier = ctx->opt.precision == HF_PREC_MIXED;
    incore = false;
    store_pending = false;
    eri_ev = ws.eri_ev;
    if (p->incore && n <= kIncoreMaxN && p->incore_max_bytes >= 0) {
        const uint64_t len = hf_pair_matrix_len(n);
        size_t free_b = 0, total_b = 0;
        HF_HIP_CHECK(ctx, hipMemGetInfo(&free_b, &total_b));
        const double want = (double)len * (mixed_tier ? 12.0 : 8.0);
        const double held = (double)ctx->d_G.n * 8.0 + (double)ctx->d_G32.n * 4.0;   // stores of an earlier SCF
        const double cap = p->incore_max_bytes > 0 ? p->incore_max_bytes : 0.5 * ((double)free_b + held);
        if (want <= cap) {
            incore = ctx->d_G.alloc_exact(len) == hipSuccess && (!mixed_tier || ctx->d_G32.alloc_exact(len) == hipSuccess);
            (void)hipGetLastError();
        }
        if (incore) {
            if (!ctx->eri_stream) HF_HIP_CHECK(ctx, hipStreamCreateWithFlags(&ctx->eri_stream, hipStreamNonBlocking));
            hipStream_t es = ctx->eri_stream;
            for (int k = 0; k < 2; ++k)
                if (!eri_ev[k]) HF_HIP_CHECK(ctx, hipEventCreate(&eri_ev[k]));
            HF_HIP_CHECK(ctx, hipEventRecord(eri_ev[0], es));
            if (int rc = hf_eri_pairs_impl(ctx, ctx->d_G.p, len, es)) return rc;
            if (mixed_tier)
                hipLaunchKernelGGL(to_f32_kernel, dim3(grid_for(len)), dim3(256), 0, es, ctx->d_G.p, ctx->d_G32.p,
                                   (size_t)len);
            HF_HIP_CHECK(ctx, hipEventRecord(eri_ev[1], es));
            store_pending = true;
        }
    }
    // h = T + V + E . M with the moment integrals about the origin (HartreeFock/HartreeFockAlgorithm.cpp:38)
    field = ctx->mol.has_field();
    if (field) HF_HIP_CHECK(ctx, st.Fb.alloc(3 * nn));
    if (int rc = hf_one_electron_device(ctx, st.S.p, st.T1.p, st.T2.p, s, field ? st.Fb.p : nullptr)) return rc;
    hipLaunchKernelGGL(add_kernel, dim3(grid_for(nn)), dim3(256), 0, s, st.T1.p, st.T2.p, st.h.p, nn);
    if (field) {
        LinComb lc{};
        lc.m = 4;
        lc.src[0] = st.h.p;
        lc.c[0] = 1.0;
        for (int k = 0; k < 3; ++k) {
            lc.src[k + 1] = st.Fb.p + k * nn;
            lc.c[k + 1] = ctx->mol.field[k];
        }
        hipLaunchKernelGGL(lincomb_kernel, dim3(grid_for(nn)), dim3(256), 0, s, lc, st.T3.p, nn);
        HF_HIP_CHECK(ctx, hipMemcpyAsync(st.h.p, st.T3.p, nn * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    // S^{-1/2}, symmetrised (:48-59)
    if (int rc = inverse_sqrt(ctx, st)) return rc;
    // occupations (HartreeFock/RestrictedHartreeFock.cpp:18-32, UnrestrictedHartreeFock.cpp:20-60)
    // ElectronsNumber() is alpha + beta when either is set, else the atoms' count (HartreeFock/Molecule.cpp:96-101)
    ne = ctx->mol.electrons();
    if (!uhf) {
        occ[0] = ne / 2;
        if (!occ[0]) occ[0] = 1;
        if (occ[0] > n) occ[0] = n;
        occ[1] = 0;
    } else {
        int a = ne / 2, b = ne - a;   // UHF takes alpha = floor(N/2), beta = the rest, unless they are set
        if (ctx->mol.alpha_set > 0 || ctx->mol.beta_set > 0) {
            a = ctx->mol.alpha_set;
            b = ctx->mol.beta_set;
        }
        if (a + b > 2 * n) a = b = n;
        if (a > n) { b += a - n; a = n; }
        if (b > n) b = n;
        occ[0] = a;
        occ[1] = b;
    }
    // iteration-0 Fock matrices (InitFockMatrix / InitFockMatrices, iter == 0), with the UHF
    // asymmetry on F_alpha(0,1) (UnrestrictedHartreeFock.cpp:217-221)
    hipLaunchKernelGGL(guess_kernel, dim3(grid_for(nn)), dim3(256), 0, s, st.S.p, st.h.p, p->init_guess, n,
                       uhf && p->add_asymmetry && n > 1 ? p->asymmetry : 0.0, st.F0.p, st.F0.p + nn);
    HF_HIP_CHECK(ctx, hipMemsetAsync(st.D.p, 0, 2 * nn * sizeof(double), s));
    for (int k = 0; k < nsp; ++k) {
        HF_HIP_CHECK(ctx, ws.diisF[k].alloc(Diis::kKeep * nc));
        HF_HIP_CHECK(ctx, ws.diisE[k].alloc(Diis::kKeep * nc));
        diis[k].F = ws.diisF[k].p;
        diis[k].E = ws.diisE[k].p;
        diis[k].nn = nn;
    }
    etolD = p->energy_tolerance > 0 ? p->energy_tolerance : 1e-11;
    etolN = p->energy_tolerance > 0 ? p->energy_tolerance : 1e-13;
    dtol = p->density_tolerance > 0 ? p->density_tolerance : 1e-10;
    itol = p->diis_tolerance > 0 ? p->diis_tolerance : 1e-10;
    // incremental Fock build state: J, K hold the J/K of Dprev
    // In-core mixed tier: the first two builds are full FP64 (the density changes most there), then
    // the J/K of each density change comes from the FP32 store while max|D - Dprev| <= 1e-3 (its
    // error scales with |D - Dprev|), FP64 otherwise; full rebuilds and the final phase are FP64.
    incore_mixed = incore && mixed_tier;
    incremental = p->incremental_fock != 0 || incore_mixed;
    first_full = incore_mixed ? 2 : 1;
    // J/K build time from events on the SCF stream, read at the step's own synchronisation
    jk_ev = ws.jk_ev;
    for (int k = 0; k < 2; ++k)
        if (!jk_ev[k]) HF_HIP_CHECK(ctx, hipEventCreate(&jk_ev[k]));
    rebuild_every = p->full_rebuild_every > 0 ? p->full_rebuild_every : 8;
    // mixed-precision tier: quartets below the context's fp32_threshold in FP32 until fp64_phase
    mixed = ctx->opt.precision == HF_PREC_MIXED;
    // in-core, every FP64 build is a full build from the FP64 store (see jk_plan), so one such step makes the Fock
    // matrix exactly FP64; direct FP32 quartets leave rounding in the DIIS history for a few steps
    final_fp64 = p->final_fp64_iterations > 0 ? p->final_fp64_iterations : incore_mixed ? 1 : 3;
    fp64_phase = !mixed || external;      // a caller's own steps (hf_scf_step) are all FP64
    useDIIS = p->use_diis != 0;
    eps_h.assign(2 * (size_t)n, 0.0);
    // HFGPU_TRACE=1 prints one line per iteration to stderr (the reference's per-step TRACE,
    // HartreeFock/RestrictedHartreeFock.cpp:149)
    const char* tenv = std::getenv("HFGPU_TRACE");
    trace = tenv && tenv[0] == '1';
    // the end of Init on this stream, for a batch whose steps run on another stream
    if (!ws.ready_ev) HF_HIP_CHECK(ctx, hipEventCreateWithFlags(&ws.ready_ev, hipEventDisableTiming));
    HF_HIP_CHECK(ctx, hipEventRecord(ws.ready_ev, s));
    return HF_OK;
}

// J/K of the current density for the coming Fock matrix (RestrictedHartreeFock::InitFockMatrix,
// HartreeFock/RestrictedHartreeFock.cpp:189-202; UnrestrictedHartreeFock::InitFockMatrices,
// HartreeFock/UnrestrictedHartreeFock.cpp:228-238): a full build, or with incremental builds (the in-core mixed
// tier, or incremental_fock) the J/K of D - Dprev added to the previous ones.
void ScfRun::jk_plan(JkPlan& pl) {
    ScfState& ss = ctx->scf_ws->st;
    pl = JkPlan{};
    pl.q0 = quartet_count();
    bool full = !incremental || force_full || full_builds < first_full || since_full >= rebuild_every;
    // In-core mixed: an increment pays only from the FP32 store (an FP64 increment streams the same store as a full
    // build and adds its rounding to the previous J/K), so a step that will not use it -- the FP64 phase, or a
    // density change above the FP32 limit -- builds in full.
    if (incore_mixed && !full && (fp64_phase || !(dmax_next >= 0 && dmax_next <= kIncoreFp32MaxDelta))) full = true;
    bool used_fp32 = !fp64_phase;
    if (incore && store_pending) {   // the store (own stream) must be complete
        pl.wait_store = true;
        store_pending = false;
    }
    pl.kind = incore ? JkPlan::kIncore : JkPlan::kDirect;
    pl.src = ss.D.p;
    if (full) {
        if (incore) {
            used_fp32 = !fp64_phase && !incore_mixed;
            pl.g32 = used_fp32 ? ctx->d_G32.p : nullptr;
        } else {
            pl.fp32 = !fp64_phase;
        }
        ++full_builds;
        since_full = 0;
        force_full = false;
    } else {
        // J/K are linear in D: add the J/K of D - Dprev, screened at thr / max|D - Dprev|.  max|D - Dprev| is
        // known from the previous step's energy phase (no read-back here).
        pl.sub = true;
        pl.src = ss.dD.p;
        pl.accumulate = true;
        const double dmax = dmax_next >= 0 ? dmax_next : 0.0;
        if (!incore || (incore_mixed && used_fp32)) last_dmax = dmax;
        // FP32 store only for small density changes: its rounding enters J/K scaled by |dD|, and
        // near-linearly-dependent (diffuse) bases keep |dD| large for many iterations
        if (incore_mixed && used_fp32 && !(dmax <= kIncoreFp32MaxDelta)) used_fp32 = false;
        if (incore) {
            pl.g32 = used_fp32 ? ctx->d_G32.p : nullptr;
        } else if (dmax > 0) {
            pl.thr = ctx->opt.schwarz_threshold / std::min(1.0, dmax);
            pl.fp32 = !fp64_phase;
        } else {
            pl.kind = JkPlan::kNone;     // no density change: J/K stay
        }
        ++incr_builds;
        ++since_full;
    }
    pl.copy_prev = incremental;
    last_incremental = !full;
    last_fp32 = used_fp32;
    if (last_fp32) {
        ++fp32_builds;
        fp64_steps = 0;
    } else {
        ++fp64_steps;
    }
}

void ScfRun::jk_done(const JkPlan& pl) {
    last_quartets = quartet_count() - pl.q0;
    quartets += last_quartets;
}

int ScfRun::step(int it, double& r) {
    step_iter = it;
    std::vector<ScfRun*> one{this};
    if (int rc = scf_step_phases(one, s, ctx->scf_ws->bs)) return rc;
    if (status != HF_OK) return status;
    r = rms;
    return HF_OK;
}

// ---- Calculate (HartreeFock/HartreeFockAlgorithm.cpp:105-192) as a state machine, one step at a time ----
// kPhDiis: the DIIS loop until convergence, cancellation or maxIterations; kPhAfterDiis: normalIterAfterDIIS
// steps without DIIS after a DIIS convergence; kPhPlain: the plain loop when DIIS did not converge.
// Convergence is accepted only on a step whose J/K build was full (not incremental) and, in the
// mixed-precision tier, all-FP64: the tier switches to FP64 once the criteria are met within a factor 100
// (or met outright) and then needs final_fp64_iterations FP64 steps.
void ScfRun::begin() {
    prev = std::numeric_limits<double>::infinity();
    ph_i = 0;
    ph = iter < p->max_iterations ? kPhDiis : kPhDone;
}

int ScfRun::after_step() {
    auto stop = [&]() { return p->terminate && __atomic_load_n(p->terminate, __ATOMIC_ACQUIRE) != 0; };
    auto accept = [&](bool met, bool nearly) {
        if (mixed && !fp64_phase && (nearly || met)) {
            fp64_phase = true;
            return false;
        }
        if (!met) return false;
        if (last_incremental) {
            force_full = true;
            return false;
        }
        return !(mixed && (last_fp32 || fp64_steps < final_fp64));
    };
    auto trace_line = [&]() {
        if (trace)
            std::fprintf(stderr, "hf_scf %s iter %4d  E % .12f  dE % .3e  |dD| %.3e  diis %.3e  jk %s max|dD| %.2e quartets %.3e\n",
                         uhf ? "UHF" : "RHF", step_iter, energy, energy - prev, rms, lastErr,
                         last_incremental ? "incr" : "full", last_dmax, last_quartets);
    };
    // a non-finite energy or density change (coincident nuclei, a corrupted density) fails the SCF at
    // once; the reference would iterate to maxIterations on NaN
    const bool finite = std::isfinite(energy) && std::isfinite(rms);
    cur = energy;
    // after the DIIS loop: normalIterAfterDIIS plain steps when DIIS converged, else the plain loop
    auto after_diis_loop = [&]() {
        if (!cancelled && useDIIS && iter < p->max_diis_iterations && converged && p->normal_iter_after_diis) {
            useDIIS = false;
            ph = kPhAfterDiis;
            ph_i = 0;
            return;
        }
        ph = (!cancelled && !converged && iter < p->max_iterations) ? kPhPlain : kPhDone;
    };
    switch (ph) {
        case kPhDiis: {
            trace_line();
            if (!finite) break;
            const double de = std::fabs(prev - cur), etol = useDIIS ? etolD : etolN;
            if (accept(de <= etol && rms < dtol && lastErr < itol,
                       de <= 100 * etol && rms < 100 * dtol && lastErr < 100 * itol)) {
                converged = true;
                after_diis_loop();
            } else if (stop()) {
                cancelled = true;
                after_diis_loop();
            } else {
                prev = cur;
                if (++iter >= p->max_iterations) after_diis_loop();
            }
            break;
        }
        case kPhAfterDiis: {
            if (!finite) break;
            prev = cur;
            const bool c = stop();
            if (c) cancelled = true;
            if (c || ++ph_i >= p->normal_iter_after_diis) {
                iter += p->normal_iter_after_diis;
                useDIIS = p->use_diis != 0;
                ph = (!cancelled && !converged && iter < p->max_iterations) ? kPhPlain : kPhDone;
            }
            break;
        }
        case kPhPlain: {
            trace_line();
            if (!finite) break;
            const double de = std::fabs(prev - cur);
            if (accept(de <= etolN && rms < dtol, de <= 100 * etolN && rms < 100 * dtol)) {
                converged = true;
                ph = kPhDone;
            } else if (stop()) {
                cancelled = true;
                ph = kPhDone;
            } else {
                prev = cur;
                if (++iter >= p->max_iterations) ph = kPhDone;
            }
            break;
        }
        default: break;
    }
    if (!finite) {
        ph = kPhDone;
        return ctx->fail(HF_E_NUMERIC, "SCF energy or density is not finite");
    }
    return HF_OK;
}

int ScfRun::calculate() {
    begin();
    while (!done()) {
        double r = 0;
        if (int rc = step(next_iter(), r)) return rc;
        if (int rc = after_step()) return rc;
    }
    return HF_OK;
}

int ScfRun::finish(hf_scf_result* r, double* D_out) {
    ScfState& st = ctx->scf_ws->st;
    // orbital energies (HOMO) of the last Fock matrices: eigenvalues only
    if (orbitals) {
        for (int k = 0; k < nsp; ++k) {
            if (last_diag[k]) continue;
            if (int rc = gemm(ctx, false, false, n, n, n, 1.0, st.Fa.p + k * nn, st.X.p, 0.0, st.T1.p)) return rc;
            if (int rc = gemm(ctx, false, false, n, n, n, 1.0, st.X.p, st.T1.p, 0.0, st.T2.p)) return rc;
            BL(rocsolver_dsyevd(ctx->blas, rocblas_evect_none, rocblas_fill_upper, n, st.T2.p, n, st.eps.p + k * n,
                                st.work.p, st.info.p));
        }
        HF_HIP_CHECK(ctx, hipMemcpyAsync(eps_h.data(), st.eps.p, nsp * n * sizeof(double), hipMemcpyDeviceToHost, s));
        HF_HIP_CHECK(ctx, hipStreamSynchronize(s));
        if (!uhf)
            homo = eps_h[occ[0] - 1];
        else
            homo = std::max(occ[0] ? eps_h[occ[0] - 1] : 0.0, occ[1] ? eps_h[n + occ[1] - 1] : 0.0);
    }
    // MP2 from the orbitals of the last Fock matrices (the reference's C and eigenvalues after
    // Calculate(), used by CalculateMp2Energy) and the in-core ERI store
    if (orbitals && p->mp2 && !cancelled) {
        for (int k = 0; k < nsp; ++k) {
            if (int rc = gemm(ctx, false, false, n, n, n, 1.0, st.Fa.p + k * nn, st.X.p, 0.0, st.T1.p)) return rc;
            if (int rc = gemm(ctx, false, false, n, n, n, 1.0, st.X.p, st.T1.p, 0.0, st.T2.p)) return rc;
            if (int rc = scf_syevd(ctx, st, ctx->stream, st.T2.p, st.eps.p + k * n)) return rc;
            if (int rc = gemm(ctx, false, false, n, n, n, 1.0, st.X.p, st.T2.p, 0.0, st.C.p + k * nn)) return rc;
        }
        if (incore && store_pending) {
            HF_HIP_CHECK(ctx, hipStreamWaitEvent(s, eri_ev[1], 0));
            store_pending = false;
        }
        if (!incore) {
            if (n > kIncoreMaxN) return ctx->fail(HF_E_UNSUPPORTED, "MP2 needs the in-core ERI store (n <= 512)");
            const uint64_t len = hf_pair_matrix_len(n);
            HF_HIP_CHECK(ctx, ctx->d_G.alloc_exact(len));
            if (int rc = hf_eri_pairs_impl(ctx, ctx->d_G.p, len, s)) return rc;
        }
        double e2 = 0;
        if (!uhf) {
            if (int rc = hf_mp2_impl(ctx, s, occ[0], st.C.p, st.eps.p, ctx->d_G.p, false, e2)) return rc;
        } else {
            for (int k = 0; k < 2; ++k) {
                double ek = 0;
                if (int rc = hf_mp2_impl(ctx, s, occ[k], st.C.p + k * nn, st.eps.p + k * n, ctx->d_G.p, true, ek))
                    return rc;
                e2 += ek;
            }
            e2 *= 0.25;   // UnrestrictedHartreeFock::CalculateMp2Energy's formula, as written
        }
        r->mp2_energy = e2;
    }
    r->energy = cur;
    r->homo = homo;
    r->iterations = iter + (converged ? 1 : 0);
    r->converged = converged ? 1 : 0;
    r->jk_ms = jk_ms;
    r->diag_steps = diag_steps;
    r->sp2_steps = sp2_total;
    r->full_builds = full_builds;
    r->incremental_builds = incr_builds;
    r->quartets = quartets;
    r->fp32_builds = fp32_builds;
    r->incore = incore ? 1 : 0;
    if (incore) {
        float ms = 0;
        HF_HIP_CHECK(ctx, hipEventSynchronize(eri_ev[1]));
        HF_HIP_CHECK(ctx, hipEventElapsedTime(&ms, eri_ev[0], eri_ev[1]));
        r->eri_ms = ms;
    }
    if (uhf && orbitals) {
        // <S^2> = ((Na-Nb)/2)^2 + (Na+Nb)/2 - sum_ij ((C_a,occ)^T S C_b,occ)_ij^2
        //       = ((Na-Nb)/2)^2 + (Na+Nb)/2 - Tr(P_a S P_b S), P = C_occ C_occ^T the last new densities
        double ov = 0;
        if (occ[0] > 0 && occ[1] > 0) {
            if (int rc = gemm(ctx, false, false, n, n, n, 1.0, st.Dn.p, st.S.p, 0.0, st.T1.p)) return rc;
            if (int rc = gemm(ctx, false, false, n, n, n, 1.0, st.S.p, st.Dn.p + nn, 0.0, st.T2.p)) return rc;
            DotJobs jobs{};
            jobs.m = 1;
            jobs.x[0] = st.T1.p;
            jobs.y[0] = st.T2.p;
            launch_dots(s, jobs, nn, st.redpart.p, st.red.p);
            if (int rc = hf_readback(ctx, s, &ov, st.red.p, 1)) return rc;
        }
        const double sz = 0.5 * (occ[0] - occ[1]);
        r->s2 = sz * sz + 0.5 * (occ[0] + occ[1]) - ov;
    }
    if (D_out) {
        HF_HIP_CHECK(ctx, hipMemcpyAsync(D_out, st.D.p, nsp * nn * sizeof(double), hipMemcpyDeviceToHost, s));
        HF_HIP_CHECK(ctx, hipStreamSynchronize(s));
    }
    r->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    if (external) return HF_OK;
    if (cancelled) return ctx->fail(HF_E_CANCELLED, "SCF cancelled");
    return converged ? HF_OK : ctx->fail(HF_E_NOT_CONVERGED, "SCF did not converge");
}

void hf_scf_session_free(hf_ctx* ctx) {
    delete ctx->scf_run;
    ctx->scf_run = nullptr;
}

extern "C" {

int hf_core_hamiltonian(hf_ctx* ctx, double* h_out) {
    if (!ctx) return HF_E_ARG;
    if (!ctx->have_mol) return ctx->fail(HF_E_STATE, "no molecule set");
    if (!h_out) return ctx->fail(HF_E_ARG, "null output");
    const int n = ctx->mol.nbf;
    const size_t nn = (size_t)n * n;
    hipStream_t s = ctx->stream;
    DevBuf<double> w;
    HF_HIP_CHECK(ctx, w.alloc(6 * nn));
    // h = T + V + E . M (HartreeFock/HartreeFockAlgorithm.cpp:38), the matrix hf_scf starts from
    if (int rc = hf_one_electron_device(ctx, w.p, w.p + nn, w.p + 2 * nn, s, w.p + 3 * nn)) return rc;
    LinComb lc{};
    lc.m = 5;
    lc.src[0] = w.p + nn;
    lc.c[0] = 1.0;
    lc.src[1] = w.p + 2 * nn;
    lc.c[1] = 1.0;
    for (int k = 0; k < 3; ++k) {
        lc.src[k + 2] = w.p + (3 + k) * nn;
        lc.c[k + 2] = ctx->mol.field[k];
    }
    hipLaunchKernelGGL(lincomb_kernel, dim3(grid_for(nn)), dim3(256), 0, s, lc, w.p, nn);
    HF_HIP_CHECK(ctx, hipGetLastError());
    HF_HIP_CHECK(ctx, hipMemcpyAsync(h_out, w.p, nn * sizeof(double), hipMemcpyDeviceToHost, s));
    HF_HIP_CHECK(ctx, hipStreamSynchronize(s));
    return HF_OK;
}

void hf_scf_defaults(hf_scf_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->unrestricted = 0;
    p->max_iterations = 3000;
    p->alpha = 0.75;
    p->init_guess = 0.75;
    p->use_diis = 1;
    p->max_diis_iterations = 1000;
    p->normal_iter_after_diis = 500;
    p->add_asymmetry = 1;
    p->asymmetry = 0.1;
    p->density_method = 0;
    p->incremental_fock = 0;
    p->full_rebuild_every = 8;
    p->final_fp64_iterations = 0;
    p->incore = 0;
    p->incore_max_bytes = 0;
    p->mp2 = 0;
}

extern "C" int hf_scf_begin(hf_ctx* ctx, const hf_scf_params* p) {
    if (!ctx) return HF_E_ARG;
    if (!ctx->have_mol) return ctx->fail(HF_E_STATE, "no molecule set");
    if (!p) return ctx->fail(HF_E_ARG, "null params");
    hf_scf_session_free(ctx);
    ctx->scf_run = new (std::nothrow) ScfRun(ctx, *p);
    if (!ctx->scf_run) return ctx->fail(HF_E_OOM, "SCF session");
    ctx->scf_run->external = true;
    ctx->scf_run->fp64_phase = true;       // a caller's own steps are all FP64
    if (int rc = ctx->scf_run->init()) {
        hf_scf_session_free(ctx);
        return rc;
    }
    return HF_OK;
}

extern "C" int hf_scf_step(hf_ctx* ctx, int iter, double* energy, double* rms, double* diis_error) {
    if (!ctx) return HF_E_ARG;
    ScfRun* run = ctx->scf_run;
    if (!run) return ctx->fail(HF_E_STATE, "no SCF session (hf_scf_begin)");
    if (iter < 0) return ctx->fail(HF_E_ARG, "negative iteration");
    double d = 0;
    if (int rc = run->step(iter, d)) return rc;
    run->cur = run->energy;
    run->iter = iter + 1;
    if (energy) *energy = run->energy;
    if (rms) *rms = d;
    if (diis_error) *diis_error = run->lastErr;
    if (!std::isfinite(run->energy) || !std::isfinite(d))
        return ctx->fail(HF_E_NUMERIC, "SCF energy or density is not finite");
    return HF_OK;
}

extern "C" int hf_scf_end(hf_ctx* ctx, hf_scf_result* r, double* D_out) {
    if (!ctx) return HF_E_ARG;
    ScfRun* run = ctx->scf_run;
    if (!run) return ctx->fail(HF_E_STATE, "no SCF session (hf_scf_begin)");
    hf_scf_result tmp;
    if (!r) r = &tmp;
    std::memset(r, 0, sizeof(*r));
    const int rc = run->finish(r, D_out);
    hf_scf_session_free(ctx);
    return rc;
}

int hf_scf(hf_ctx* ctx, const hf_scf_params* p, hf_scf_result* r, double* D_out) {
    if (!ctx) return HF_E_ARG;
    if (!ctx->have_mol) return ctx->fail(HF_E_STATE, "no molecule set");
    if (!p || !r) return ctx->fail(HF_E_ARG, "null params/result");
    std::memset(r, 0, sizeof(*r));
    ScfRun run(ctx, *p);
    if (int rc = run.init()) return rc;
    if (int rc = run.calculate()) return rc;
    return run.finish(r, D_out);
}

}  // extern "C"
