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
// All matrices live in HBM; the dense steps use rocBLAS dgemm and rocSOLVER dsyevd
// (FP64).  Only eigenvalues, a few scalars and the DIIS inner products come back to
// the host each iteration.  Matrices are symmetric except the MO coefficients C,
// which are stored column-major (column j = orbital j) as rocSOLVER returns them.
#include <rocsolver/rocsolver.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <deque>
#include <limits>

#include "hf_internal.h"

namespace {

__global__ void fock_kernel(const double* __restrict__ h, const double* __restrict__ J, const double* __restrict__ K,
                            double sk, double* __restrict__ F, int n, int symmetrize) {
    const size_t nn = (size_t)n * n;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nn; idx += (size_t)gridDim.x * blockDim.x) {
        const size_t i = idx / n, j = idx % n, t = j * n + i;
        const double a = h[idx] + J[idx] - sk * K[idx];
        F[idx] = symmetrize ? 0.5 * (a + (h[t] + J[t] - sk * K[t])) : a;
    }
}

__global__ void add_kernel(const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ c,
                           size_t nn) {
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nn; idx += (size_t)gridDim.x * blockDim.x)
        c[idx] = a[idx] + b[idx];
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

__global__ void antisym_kernel(const double* __restrict__ Y, double* __restrict__ E, int n) {
    const size_t nn = (size_t)n * n;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nn; idx += (size_t)gridDim.x * blockDim.x) {
        const size_t i = idx % n, j = idx / n;   // column-major (i row, j column)
        E[idx] = Y[idx] - Y[i * n + j];
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

__global__ void mix_kernel(const double* __restrict__ Dn, double* __restrict__ D, double a, size_t nn, int replace) {
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nn; idx += (size_t)gridDim.x * blockDim.x)
        D[idx] = replace ? Dn[idx] : a * Dn[idx] + (1.0 - a) * D[idx];
}

// Deterministic single-block reductions: out[r] = sum_i x_r[i] * y_r[i]  or  sum_i (x_r[i]-y_r[i])^2
struct DotJobs {
    const double* x[8];
    const double* y[8];
    int diff[8];
    int m;
};
__global__ void __launch_bounds__(1024) dots_kernel(DotJobs jobs, size_t nn, double* __restrict__ out) {
    __shared__ double part[1024];
    for (int r = 0; r < jobs.m; ++r) {
        double acc = 0;
        const double* __restrict__ x = jobs.x[r];
        const double* __restrict__ y = jobs.y[r];
        if (jobs.diff[r]) {
            for (size_t i = threadIdx.x; i < nn; i += blockDim.x) {
                const double d = x[i] - y[i];
                acc += d * d;
            }
        } else {
            for (size_t i = threadIdx.x; i < nn; i += blockDim.x) acc += x[i] * y[i];
        }
        part[threadIdx.x] = acc;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[r] = part[0];
        __syncthreads();
    }
}

int grid_for(size_t n) {
    size_t b = (n + 255) / 256;
    return (int)std::min<size_t>(std::max<size_t>(b, 1), 4096);
}

// Column-pivoted Householder QR solve of the (m x m) DIIS system (row-major A).
std::vector<double> qr_solve(int m, std::vector<double> A, std::vector<double> b) {
    std::vector<int> perm(m);
    for (int i = 0; i < m; ++i) perm[i] = i;
    double maxnorm = 0;
    for (int j = 0; j < m; ++j) {
        double s = 0;
        for (int i = 0; i < m; ++i) s += A[i * m + j] * A[i * m + j];
        maxnorm = std::max(maxnorm, std::sqrt(s));
    }
    int rank = m;
    for (int k = 0; k < m; ++k) {
        int piv = k;
        double best = -1;
        for (int j = k; j < m; ++j) {
            double s = 0;
            for (int i = k; i < m; ++i) s += A[i * m + j] * A[i * m + j];
            if (s > best) { best = s; piv = j; }
        }
        if (piv != k) {
            for (int i = 0; i < m; ++i) std::swap(A[i * m + k], A[i * m + piv]);
            std::swap(perm[k], perm[piv]);
        }
        double nrm = 0;
        for (int i = k; i < m; ++i) nrm += A[i * m + k] * A[i * m + k];
        nrm = std::sqrt(nrm);
        if (nrm <= std::numeric_limits<double>::epsilon() * maxnorm * m) { rank = k; break; }
        const double alpha = A[k * m + k] > 0 ? -nrm : nrm;
        std::vector<double> v(m, 0.0);
        for (int i = k; i < m; ++i) v[i] = A[i * m + k];
        v[k] -= alpha;
        double vv = 0;
        for (int i = k; i < m; ++i) vv += v[i] * v[i];
        if (vv == 0) continue;
        for (int j = k; j < m; ++j) {
            double s = 0;
            for (int i = k; i < m; ++i) s += v[i] * A[i * m + j];
            s *= 2 / vv;
            for (int i = k; i < m; ++i) A[i * m + j] -= s * v[i];
        }
        double s = 0;
        for (int i = k; i < m; ++i) s += v[i] * b[i];
        s *= 2 / vv;
        for (int i = k; i < m; ++i) b[i] -= s * v[i];
    }
    std::vector<double> y(m, 0.0), x(m, 0.0);
    for (int i = rank - 1; i >= 0; --i) {
        double s = b[i];
        for (int j = i + 1; j < rank; ++j) s -= A[i * m + j] * y[j];
        y[i] = s / A[i * m + i];
    }
    for (int i = 0; i < m; ++i) x[perm[i]] = y[i];
    return x;
}

// DIIS history of one spin: device slots for F and e, host inner products.
struct Diis {
    static constexpr int kKeep = 6, kFirst = 5;
    DevBuf<double> F, E;          // kKeep * nn each
    std::deque<int> order;        // slot indices, oldest first
    double B[kKeep][kKeep] = {};  // B[slot][slot]
    size_t nn = 0;
    int next_slot() {
        if ((int)order.size() < kKeep) return (int)order.size();
        return order.front();
    }
};

struct ScfState {
    int n = 0;
    size_t nn = 0;
    DevBuf<double> S, h, X, F0, Fa, Fb, J, K, D, Dn, C, eps, T1, T2, err, work, red;
    DevBuf<rocblas_int> info;
};

#define BL(call)                                                                                        \
    do {                                                                                                \
        rocblas_status s_ = (call);                                                                     \
        if (s_ != rocblas_status_success)                                                               \
            return ctx->fail(HF_E_HIP, std::string(#call) + ": rocblas status " + std::to_string((int)s_)); \
    } while (0)

int gemm(hf_ctx* ctx, bool ta, bool tb, int m, int n, int k, double alpha, const double* A, const double* B,
         double beta, double* C) {
    BL(rocblas_dgemm(ctx->blas, ta ? rocblas_operation_transpose : rocblas_operation_none,
                     tb ? rocblas_operation_transpose : rocblas_operation_none, m, n, k, &alpha, A, ta ? k : m, B,
                     tb ? n : k, &beta, C, m));
    return HF_OK;
}

// Symmetric eigenproblem, in place: A -> eigenvectors (column-major), w ascending.
// HFGPU_EIGEN selects the rocSOLVER routine: syevd (divide and conquer, default), syevj (Jacobi),
// syevdj, syev (QR iteration).  All FP64.
int syevd(hf_ctx* ctx, ScfState& st, double* A, double* w) {
    static const char* env = std::getenv("HFGPU_EIGEN");
    static const int alg = !env ? 0 : (!std::strcmp(env, "syevj") ? 1 : !std::strcmp(env, "syevdj") ? 2
                                                                     : !std::strcmp(env, "syev") ? 3 : 0);
    if (alg == 1) {
        BL(rocsolver_dsyevj(ctx->blas, rocblas_esort_ascending, rocblas_evect_original, rocblas_fill_upper, st.n, A,
                            st.n, 0.0, st.red.p + 8, 100, st.info.p + 1, w, st.info.p));
    } else if (alg == 2) {
        BL(rocsolver_dsyevdj(ctx->blas, rocblas_evect_original, rocblas_fill_upper, st.n, A, st.n, w, st.info.p));
    } else if (alg == 3) {
        BL(rocsolver_dsyev(ctx->blas, rocblas_evect_original, rocblas_fill_upper, st.n, A, st.n, w, st.work.p,
                           st.info.p));
    } else {
        BL(rocsolver_dsyevd(ctx->blas, rocblas_evect_original, rocblas_fill_upper, st.n, A, st.n, w, st.work.p,
                            st.info.p));
    }
    return HF_OK;
}

// Add (F, e) to a DIIS history; returns true when an extrapolation is possible.
int diis_add(hf_ctx* ctx, ScfState& st, Diis& d, const double* Fsrc, const double* esrc, bool& ready) {
    hipStream_t s = ctx->stream;
    const int slot = d.next_slot();
    if ((int)d.order.size() == Diis::kKeep) d.order.pop_front();
    HF_HIP_CHECK(ctx, hipMemcpyAsync(d.F.p + slot * st.nn, Fsrc, st.nn * sizeof(double), hipMemcpyDeviceToDevice, s));
    HF_HIP_CHECK(ctx, hipMemcpyAsync(d.E.p + slot * st.nn, esrc, st.nn * sizeof(double), hipMemcpyDeviceToDevice, s));
    d.order.push_back(slot);
    DotJobs jobs{};
    jobs.m = (int)d.order.size();
    for (int r = 0; r < jobs.m; ++r) {
        jobs.x[r] = d.E.p + slot * st.nn;
        jobs.y[r] = d.E.p + d.order[r] * st.nn;
        jobs.diff[r] = 0;
    }
    hipLaunchKernelGGL(dots_kernel, dim3(1), dim3(1024), 0, s, jobs, st.nn, st.red.p);
    double h[8];
    HF_HIP_CHECK(ctx, hipMemcpyAsync(h, st.red.p, jobs.m * sizeof(double), hipMemcpyDeviceToHost, s));
    HF_HIP_CHECK(ctx, hipStreamSynchronize(s));
    for (int r = 0; r < jobs.m; ++r) d.B[slot][d.order[r]] = d.B[d.order[r]][slot] = h[r];
    ready = (int)d.order.size() >= Diis::kFirst;
    return HF_OK;
}

// Extrapolate F = sum c_i F_i; returns the error estimate sqrt(B_last,last + B_prev,prev).
int diis_estimate(hf_ctx* ctx, ScfState& st, Diis& d, double* Fout, double& est) {
    const int m = (int)d.order.size(), mp = m + 1;
    std::vector<double> B((size_t)mp * mp, 0.0), rhs(mp, 0.0);
    est = 0;
    for (int i = 0; i < m; ++i) {
        for (int j = 0; j < m; ++j) B[i * mp + j] = d.B[d.order[i]][d.order[j]];
        if (i == m - 1 || i == m - 2) est += B[i * mp + i];
        B[m * mp + i] = B[i * mp + m] = 1;
    }
    est = std::sqrt(est);
    rhs[m] = 1;
    std::vector<double> c = qr_solve(mp, B, rhs);
    LinComb lc{};
    lc.m = m;
    for (int i = 0; i < m; ++i) {
        lc.src[i] = d.F.p + d.order[i] * st.nn;
        lc.c[i] = c[i];
    }
    hipLaunchKernelGGL(lincomb_kernel, dim3(grid_for(st.nn)), dim3(256), 0, ctx->stream, lc, Fout, st.nn);
    HF_HIP_CHECK(ctx, hipGetLastError());
    return HF_OK;
}

}  // namespace

extern "C" {

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
}

int hf_scf(hf_ctx* ctx, const hf_scf_params* p, hf_scf_result* r, double* D_out) {
    if (!ctx) return HF_E_ARG;
    if (!ctx->have_mol) return ctx->fail(HF_E_STATE, "no molecule set");
    if (!p || !r) return ctx->fail(HF_E_ARG, "null params/result");
    const auto t_start = std::chrono::steady_clock::now();
    std::memset(r, 0, sizeof(*r));
    hipStream_t s = ctx->stream;
    const bool uhf = p->unrestricted != 0;
    ScfState st;
    const int n = ctx->mol.nbf;
    st.n = n;
    st.nn = (size_t)n * n;
    const size_t nn = st.nn;
    // ---- Init (HartreeFock/HartreeFockAlgorithm.cpp:21-101) ----
    const double Enuc = hfhost::nuclear_repulsion(ctx->mol);
    const int nsp = uhf ? 2 : 1;
    HF_HIP_CHECK(ctx, st.S.alloc(nn));
    HF_HIP_CHECK(ctx, st.h.alloc(nn));
    HF_HIP_CHECK(ctx, st.X.alloc(nn));
    HF_HIP_CHECK(ctx, st.F0.alloc(2 * nn));
    HF_HIP_CHECK(ctx, st.Fa.alloc(2 * nn));
    HF_HIP_CHECK(ctx, st.J.alloc(nn));
    HF_HIP_CHECK(ctx, st.K.alloc(2 * nn));
    HF_HIP_CHECK(ctx, st.D.alloc(2 * nn));
    HF_HIP_CHECK(ctx, st.Dn.alloc(2 * nn));
    HF_HIP_CHECK(ctx, st.C.alloc(2 * nn));
    HF_HIP_CHECK(ctx, st.eps.alloc(2 * n));
    HF_HIP_CHECK(ctx, st.T1.alloc(nn));
    HF_HIP_CHECK(ctx, st.T2.alloc(nn));
    HF_HIP_CHECK(ctx, st.err.alloc(nn));
    HF_HIP_CHECK(ctx, st.work.alloc(n));
    HF_HIP_CHECK(ctx, st.red.alloc(16));
    HF_HIP_CHECK(ctx, st.info.alloc(2));
    // S, T, V on the device; h = T + V
    if (int rc = hf_one_electron_device(ctx, st.S.p, st.T1.p, st.T2.p, s)) return rc;
    hipLaunchKernelGGL(add_kernel, dim3(grid_for(nn)), dim3(256), 0, s, st.T1.p, st.T2.p, st.h.p, nn);
    // S^{-1/2} = U s^{-1/2} U^T, symmetrised (:48-59)
    HF_HIP_CHECK(ctx, hipMemcpyAsync(st.T1.p, st.S.p, nn * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (int rc = syevd(ctx, st, st.T1.p, st.eps.p)) return rc;
    hipLaunchKernelGGL(scale_cols_kernel, dim3(grid_for(nn)), dim3(256), 0, s, st.T1.p, st.eps.p, st.T2.p, n);
    if (int rc = gemm(ctx, false, true, n, n, n, 1.0, st.T2.p, st.T1.p, 0.0, st.X.p)) return rc;
    hipLaunchKernelGGL(sym_kernel, dim3(grid_for(nn)), dim3(256), 0, s, st.X.p, n);
    // occupations (HartreeFock/RestrictedHartreeFock.cpp:18-32, UnrestrictedHartreeFock.cpp:20-60)
    const int ne = ctx->mol.electrons();
    int occ[2];
    if (!uhf) {
        occ[0] = ne / 2;
        if (!occ[0]) occ[0] = 1;
        if (occ[0] > n) occ[0] = n;
        occ[1] = 0;
    } else {
        int a = ne / 2, b = ne - a;   // Molecule::Init split; UHF then takes alpha=ne/2, beta=rest
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
    Diis diis[2];
    for (int k = 0; k < nsp; ++k) {
        HF_HIP_CHECK(ctx, diis[k].F.alloc(Diis::kKeep * nn));
        HF_HIP_CHECK(ctx, diis[k].E.alloc(Diis::kKeep * nn));
        diis[k].nn = nn;
    }
    const double etolD = p->energy_tolerance > 0 ? p->energy_tolerance : 1e-11;
    const double etolN = p->energy_tolerance > 0 ? p->energy_tolerance : 1e-13;
    const double dtol = p->density_tolerance > 0 ? p->density_tolerance : 1e-10;
    const double itol = p->diis_tolerance > 0 ? p->diis_tolerance : 1e-10;
    double energy = 0, homo = 0, lastErr = 0, jk_ms = 0;
    bool useDIIS = p->use_diis != 0;
    std::vector<double> eps_h(2 * (size_t)n);

    // ---- one Step (RHF::Step / UHF::Step) ----
    auto step = [&](int iter, double& rms) -> int {
        double* F[2] = {st.Fa.p, st.Fa.p + nn};
        if (iter == 0) {
            for (int k = 0; k < nsp; ++k)
                HF_HIP_CHECK(ctx, hipMemcpyAsync(F[k], st.F0.p + k * nn, nn * sizeof(double), hipMemcpyDeviceToDevice, s));
        } else {
            const auto t0 = std::chrono::steady_clock::now();
            if (int rc = hf_jk_device_impl(ctx, nsp, st.D.p, st.J.p, st.K.p, s)) return rc;
            HF_HIP_CHECK(ctx, hipStreamSynchronize(s));
            jk_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            for (int k = 0; k < nsp; ++k)
                hipLaunchKernelGGL(fock_kernel, dim3(grid_for(nn)), dim3(256), 0, s, st.h.p, st.J.p, st.K.p + k * nn,
                                   uhf ? 1.0 : 0.5, F[k], n, uhf ? 0 : 1);
        }
        // DIISStep: e = S D F - F D S
        bool used = false;
        if (useDIIS && iter && iter < p->max_diis_iterations) {
            bool ready[2] = {false, false};
            for (int k = 0; k < nsp; ++k) {
                if (int rc = gemm(ctx, false, false, n, n, n, 1.0, st.S.p, st.D.p + k * nn, 0.0, st.T1.p)) return rc;
                if (int rc = gemm(ctx, false, false, n, n, n, 1.0, st.T1.p, F[k], 0.0, st.T2.p)) return rc;
                hipLaunchKernelGGL(antisym_kernel, dim3(grid_for(nn)), dim3(256), 0, s, st.T2.p, st.err.p, n);
                if (int rc = diis_add(ctx, st, diis[k], F[k], st.err.p, ready[k])) return rc;
            }
            if (ready[nsp - 1]) {
                double e0 = 0, e1 = 0;
                if (int rc = diis_estimate(ctx, st, diis[nsp - 1], F[nsp - 1], e1)) return rc;
                if (uhf)
                    if (int rc = diis_estimate(ctx, st, diis[0], F[0], e0)) return rc;
                lastErr = e1 + e0;
                used = true;
            }
        } else
            lastErr = 0;
        // orthogonalise, diagonalise, back-transform, new density
        for (int k = 0; k < nsp; ++k) {
            if (int rc = gemm(ctx, false, false, n, n, n, 1.0, F[k], st.X.p, 0.0, st.T1.p)) return rc;
            if (int rc = gemm(ctx, false, false, n, n, n, 1.0, st.X.p, st.T1.p, 0.0, st.T2.p)) return rc;
            if (int rc = syevd(ctx, st, st.T2.p, st.eps.p + k * n)) return rc;
            if (int rc = gemm(ctx, false, false, n, n, n, 1.0, st.X.p, st.T2.p, 0.0, st.C.p + k * nn)) return rc;
            if (int rc = gemm(ctx, false, true, n, n, occ[k], uhf ? 1.0 : 2.0, st.C.p + k * nn, st.C.p + k * nn, 0.0,
                              st.Dn.p + k * nn))
                return rc;
        }
        // energy pieces and density change
        DotJobs jobs{};
        jobs.m = 0;
        for (int k = 0; k < nsp; ++k) {
            jobs.x[jobs.m] = st.Dn.p + k * nn; jobs.y[jobs.m] = st.h.p; jobs.diff[jobs.m++] = 0;
            jobs.x[jobs.m] = st.Dn.p + k * nn; jobs.y[jobs.m] = st.D.p + k * nn; jobs.diff[jobs.m++] = 1;
        }
        hipLaunchKernelGGL(dots_kernel, dim3(1), dim3(1024), 0, s, jobs, nn, st.red.p);
        double red[8];
        HF_HIP_CHECK(ctx, hipMemcpyAsync(red, st.red.p, jobs.m * sizeof(double), hipMemcpyDeviceToHost, s));
        HF_HIP_CHECK(ctx, hipMemcpyAsync(eps_h.data(), st.eps.p, nsp * n * sizeof(double), hipMemcpyDeviceToHost, s));
        HF_HIP_CHECK(ctx, hipStreamSynchronize(s));
        double eh = 0, rms2 = 0;
        for (int k = 0; k < nsp; ++k) {
            eh += red[2 * k];
            rms2 += red[2 * k + 1];
        }
        double e = 0;
        if (!uhf) {
            e = 0.5 * eh;
            for (int v = 0; v < occ[0]; ++v) e += eps_h[v];
            homo = eps_h[occ[0] - 1];
        } else {
            e = eh;
            for (int v = 0; v < occ[0]; ++v) e += eps_h[v];
            for (int v = 0; v < occ[1]; ++v) e += eps_h[n + v];
            homo = std::max(occ[0] ? eps_h[occ[0] - 1] : 0.0, occ[1] ? eps_h[n + occ[1] - 1] : 0.0);
            e *= 0.5;
        }
        energy = e + Enuc;
        rms = std::sqrt(rms2);
        hipLaunchKernelGGL(mix_kernel, dim3(grid_for(nsp * nn)), dim3(256), 0, s, st.Dn.p, st.D.p, p->alpha,
                           nsp * nn, used ? 1 : 0);
        HF_HIP_CHECK(ctx, hipGetLastError());
        return HF_OK;
    };

    // ---- Calculate (HartreeFock/HartreeFockAlgorithm.cpp:105-192) ----
    // HFGPU_TRACE=1 prints one line per iteration to stderr (the reference's per-step TRACE,
    // HartreeFock/RestrictedHartreeFock.cpp:149).
    const char* tenv = std::getenv("HFGPU_TRACE");
    const bool trace = tenv && tenv[0] == '1';
    double cur = 0, prev = std::numeric_limits<double>::infinity(), rms = 0;
    auto trace_line = [&](int it, double rmsv) {
        if (trace)
            std::fprintf(stderr, "hf_scf %s iter %4d  E % .12f  dE % .3e  |dD| %.3e  diis %.3e\n", uhf ? "UHF" : "RHF",
                         it, energy, energy - prev, rmsv, lastErr);
    };
    int iter = 0;
    bool converged = false, cancelled = false;
    auto stop = [&]() { return p->terminate && *p->terminate; };
    for (; iter < p->max_iterations; ++iter) {
        if (int rc = step(iter, rms)) return rc;
        cur = energy;
        trace_line(iter, rms);
        if (std::fabs(prev - cur) <= (useDIIS ? etolD : etolN) && rms < dtol && lastErr < itol) {
            converged = true;
            break;
        }
        if (stop()) { cancelled = true; break; }
        prev = cur;
    }
    if (!cancelled && useDIIS && iter < p->max_diis_iterations && converged && p->normal_iter_after_diis) {
        useDIIS = false;
        for (int i = 0; i < p->normal_iter_after_diis; ++i) {
            if (int rc = step(iter + i, rms)) return rc;
            cur = energy;
            prev = cur;
            if (stop()) { cancelled = true; break; }
        }
        iter += p->normal_iter_after_diis;
        useDIIS = p->use_diis != 0;
    }
    if (!cancelled && !converged) {
        for (; iter < p->max_iterations; ++iter) {
            if (int rc = step(iter, rms)) return rc;
            cur = energy;
            trace_line(iter, rms);
            if (std::fabs(prev - cur) <= etolN && rms < dtol) { converged = true; break; }
            if (stop()) { cancelled = true; break; }
            prev = cur;
        }
    }
    r->energy = cur;
    r->homo = homo;
    r->iterations = iter + (converged ? 1 : 0);
    r->converged = converged ? 1 : 0;
    r->jk_ms = jk_ms;
    if (uhf) {
        // <S^2> = ((Na-Nb)/2)^2 + (Na+Nb)/2 - sum_ij ((C_a,occ)^T S C_b,occ)_ij^2
        if (occ[0] > 0 && occ[1] > 0) {
            if (int rc = gemm(ctx, false, false, n, occ[1], n, 1.0, st.S.p, st.C.p + nn, 0.0, st.T1.p)) return rc;
            if (int rc = gemm(ctx, true, false, occ[0], occ[1], n, 1.0, st.C.p, st.T1.p, 0.0, st.T2.p)) return rc;
        }
        std::vector<double> M((size_t)occ[0] * occ[1]);
        if (!M.empty()) {
            HF_HIP_CHECK(ctx, hipMemcpyAsync(M.data(), st.T2.p, M.size() * sizeof(double), hipMemcpyDeviceToHost, s));
            HF_HIP_CHECK(ctx, hipStreamSynchronize(s));
        }
        double ov = 0;
        for (double x : M) ov += x * x;
        const double sz = 0.5 * (occ[0] - occ[1]);
        r->s2 = sz * sz + 0.5 * (occ[0] + occ[1]) - ov;
    }
    if (D_out) {
        HF_HIP_CHECK(ctx, hipMemcpyAsync(D_out, st.D.p, nsp * nn * sizeof(double), hipMemcpyDeviceToHost, s));
        HF_HIP_CHECK(ctx, hipStreamSynchronize(s));
    }
    r->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    if (cancelled) return ctx->fail(HF_E_CANCELLED, "SCF cancelled");
    return converged ? HF_OK : ctx->fail(HF_E_NOT_CONVERGED, "SCF did not converge");
}

}  // extern "C"
