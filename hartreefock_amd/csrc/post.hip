// After the SCF: MP2 correlation energy from the in-core ERI store, and the one-electron
// properties (Mulliken charges, dipole moment).
//
// MP2 (RestrictedHartreeFock::CalculateMp2Energy, HartreeFock/RestrictedHartreeFock.cpp:251-288, and
// UnrestrictedHartreeFock::CalculateMp2Energy, HartreeFock/UnrestrictedHartreeFock.cpp:295-375).  The
// reference transforms ERIs to the MO basis one quarter at a time with hash-map caches
// (HartreeFock/IntegralsRepository.h:190-291).  Here the transform is two FP64 GEMMs on the matrix
// cores over the in-core store G (symmetric over canonical AO pairs P = (m >= n), padded positions):
//   C2[P][(i,a)] = C_mi C_na + C_ni C_ma (m > n), C_mi C_ma (m = n), 0 at padding positions
//   (ia|jb) = sum_{P,Q} C2[P][(ia)] G[P][Q] C2[Q][(jb)]  =  (C2^T G C2)[(ia)][(jb)]
// 2 ov np^2 + 2 (ov)^2 np flops (o occupied, v virtual orbitals, np ~ n(n+1)/2 + n/2), then one pass
// over the (ov)^2 MO integrals for the energy.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <vector>

#include "hf_internal.h"

namespace {

// C2 column (ia) = i * nv + a, row R = pair_pad_index(m, q) of the in-core store (padding rows 0);
// column-major with leading dimension np (the store's padded pair count).  C is column-major:
// orbital p's coefficients at C[p * n + m].
__global__ void mp2_c2_kernel(const double* __restrict__ C, int n, int nocc, int nv, int64_t np,
                              double* __restrict__ C2) {
    const int64_t R = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (R >= np) return;
    int m = (int)sqrt(2.0 * (double)R) - 1;
    if (m < 0) m = 0;
    while (hfgpu::pair_pad_base(m + 1) <= R) ++m;
    while (m > 0 && hfgpu::pair_pad_base(m) > R) --m;
    const int q = (int)(R - hfgpu::pair_pad_base(m));
    for (int ia = blockIdx.y; ia < nocc * nv; ia += gridDim.y) {
        double v = 0.0;
        if (q <= m) {
            const int i = ia / nv, a = nocc + ia % nv;
            const double* ci = C + (size_t)i * n;
            const double* ca = C + (size_t)a * n;
            v = ci[m] * ca[q];
            if (m != q) v += ci[q] * ca[m];
        }
        C2[(size_t)ia * np + R] = v;
    }
}

// Partial sums of the MP2 energy, one per block: MO[(jb) * ov + (ia)] = (ia|jb).
// same_spin: (ia|jb)^2 / d (the reference's UHF formula); else (ia|jb)(2 (ia|jb) - (ib|ja)) / d.
__global__ void __launch_bounds__(256) mp2_energy_kernel(const double* __restrict__ MO, const double* __restrict__ eps,
                                                         int nocc, int nv, int same_spin, double* __restrict__ part) {
    const int64_t ov = (int64_t)nocc * nv, total = ov * ov;
    double acc = 0.0;
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ia = x % ov, jb = x / ov;
        const int i = (int)(ia / nv), a = (int)(ia % nv), j = (int)(jb / nv), b = (int)(jb % nv);
        const double v = MO[x];
        const double d = eps[i] + eps[j] - eps[nocc + a] - eps[nocc + b];
        if (same_spin) {
            acc += v * v / d;
        } else {
            const double w = MO[((int64_t)j * nv + a) * ov + (int64_t)i * nv + b];   // (ib|ja)
            acc += v * (2.0 * v - w) / d;
        }
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Mulliken populations and dipole sums, one block each: block b < natoms sums (D S)_mm over the
// functions [bf0[b], bf0[b+1]) of atom b; block natoms + d sums D_mn M_d,mn.
__global__ void __launch_bounds__(256) props_kernel(const double* __restrict__ D, const double* __restrict__ S,
                                                    const double* __restrict__ M, const int* __restrict__ bf0,
                                                    int natoms, int n, double* __restrict__ out) {
    const int b = blockIdx.x;
    const size_t nn = (size_t)n * n;
    double acc = 0.0;
    if (b < natoms) {
        const int f0 = bf0[b], f1 = bf0[b + 1];
        const size_t cnt = (size_t)(f1 - f0) * n;
        for (size_t x = threadIdx.x; x < cnt; x += blockDim.x) {
            const int m = f0 + (int)(x / n), q = (int)(x % n);
            acc += D[(size_t)m * n + q] * S[(size_t)q * n + m];
        }
    } else {
        const double* Md = M + (size_t)(b - natoms) * nn;
        for (size_t x = threadIdx.x; x < nn; x += blockDim.x) acc += D[x] * Md[x];
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[b] = red[0];
}

__global__ void add2_kernel(const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ c, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        c[i] = a[i] + b[i];
}

#define HF_BL(call)                                                                                     \
    do {                                                                                                \
        rocblas_status s_ = (call);                                                                     \
        if (s_ != rocblas_status_success)                                                               \
            return ctx->fail(HF_E_HIP, std::string(#call) + ": rocblas status " + std::to_string((int)s_)); \
    } while (0)

}  // namespace

int hf_mp2_impl(hf_ctx* ctx, hipStream_t st, int nocc, const double* C, const double* eps, const double* G,
                bool same_spin, double& e2) {
    const int n = ctx->mol.nbf;
    const int nv = n - nocc;
    e2 = 0.0;
    if (nocc <= 0 || nv <= 0) return HF_OK;
    const int64_t np = hfgpu::pair_pad_base(n), ov = (int64_t)nocc * nv;     // the store is np x np
    DevBuf<double> C2, T1, MO, part;
    HF_HIP_CHECK(ctx, C2.alloc((size_t)(np * ov)));
    HF_HIP_CHECK(ctx, T1.alloc((size_t)(np * ov)));
    HF_HIP_CHECK(ctx, MO.alloc((size_t)(ov * ov)));
    constexpr int kParts = 1024;
    HF_HIP_CHECK(ctx, part.alloc(kParts));
    hipLaunchKernelGGL(mp2_c2_kernel, dim3((unsigned)((np + 255) / 256), (unsigned)std::min<int64_t>(ov, 65535)),
                       dim3(256), 0, st, C, n, nocc, nv, np, C2.p);
    HF_HIP_CHECK(ctx, hipGetLastError());
    HF_BL(rocblas_set_stream(ctx->blas, st));   // restored to the context stream below
    const double one = 1.0, zero = 0.0;
    // T1 (ov x np) = C2^T G
    HF_BL(rocblas_dgemm(ctx->blas, rocblas_operation_transpose, rocblas_operation_none, (rocblas_int)ov,
                        (rocblas_int)np, (rocblas_int)np, &one, C2.p, (rocblas_int)np, G, (rocblas_int)np, &zero, T1.p,
                        (rocblas_int)ov));
    // MO (ov x ov) = T1 C2
    HF_BL(rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_none, (rocblas_int)ov, (rocblas_int)ov,
                        (rocblas_int)np, &one, T1.p, (rocblas_int)ov, C2.p, (rocblas_int)np, &zero, MO.p,
                        (rocblas_int)ov));
    HF_BL(rocblas_set_stream(ctx->blas, ctx->stream));
    hipLaunchKernelGGL(mp2_energy_kernel, dim3(kParts), dim3(256), 0, st, MO.p, eps, nocc, nv, same_spin ? 1 : 0,
                       part.p);
    HF_HIP_CHECK(ctx, hipGetLastError());
    std::vector<double> h(kParts);
    HF_HIP_CHECK(ctx, hipMemcpyAsync(h.data(), part.p, kParts * sizeof(double), hipMemcpyDeviceToHost, st));
    HF_HIP_CHECK(ctx, hipStreamSynchronize(st));
    for (double x : h) e2 += x;       // fixed order: deterministic
    return HF_OK;
}

int hf_properties_impl(hf_ctx* ctx, int ndens, const double* d_D, double* charges, double* dipole, hipStream_t st) {
    const hfhost::Molecule& mol = ctx->mol;
    const int n = mol.nbf, natoms = (int)mol.Z.size();
    const size_t nn = (size_t)n * n;
    DevBuf<double> S, T, V, M, Dt, out;
    DevBuf<int> bf0;
    HF_HIP_CHECK(ctx, S.alloc(nn));
    HF_HIP_CHECK(ctx, T.alloc(nn));
    HF_HIP_CHECK(ctx, V.alloc(nn));
    HF_HIP_CHECK(ctx, M.alloc(3 * nn));
    HF_HIP_CHECK(ctx, Dt.alloc(nn));
    HF_HIP_CHECK(ctx, out.alloc(natoms + 3));
    HF_HIP_CHECK(ctx, bf0.alloc(natoms + 1));
    if (int rc = hf_one_electron_device(ctx, S.p, T.p, V.p, st, M.p)) return rc;
    // basis functions of atom a: [bf0[a], bf0[a+1]) (atom -> shell -> component order)
    std::vector<int> first(natoms + 1, n);
    for (const auto& sh : mol.shells) first[sh.atom] = std::min(first[sh.atom], sh.bf);
    for (int a = natoms - 1; a >= 0; --a) first[a] = std::min(first[a], first[a + 1]);
    HF_HIP_CHECK(ctx, hipMemcpyAsync(bf0.p, first.data(), (natoms + 1) * sizeof(int), hipMemcpyHostToDevice, st));
    const double* Dtot = d_D;
    if (ndens == 2) {
        hipLaunchKernelGGL(add2_kernel, dim3(256), dim3(256), 0, st, d_D, d_D + nn, Dt.p, nn);
        Dtot = Dt.p;
    }
    hipLaunchKernelGGL(props_kernel, dim3(natoms + 3), dim3(256), 0, st, Dtot, S.p, M.p, bf0.p, natoms, n, out.p);
    HF_HIP_CHECK(ctx, hipGetLastError());
    std::vector<double> h(natoms + 3);
    HF_HIP_CHECK(ctx, hipMemcpyAsync(h.data(), out.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HF_HIP_CHECK(ctx, hipStreamSynchronize(st));
    for (int a = 0; a < natoms; ++a)
        if (charges) charges[a] = mol.Z[a] - h[a];
    if (dipole) {
        for (int d = 0; d < 3; ++d) {
            double nuc = 0;
            for (int a = 0; a < natoms; ++a) nuc += mol.Z[a] * mol.xyz[3 * a + d];
            dipole[d] = nuc - h[natoms + d];
        }
    }
    return HF_OK;
}

extern "C" int hf_properties(hf_ctx* ctx, int ndens, const double* D, double* charges, double* dipole) {
    if (!ctx) return HF_E_ARG;
    if (!ctx->have_mol) return ctx->fail(HF_E_STATE, "no molecule set");
    if (!D || ndens < 1 || ndens > 2) return ctx->fail(HF_E_ARG, "bad arguments");
    const size_t nn = (size_t)ctx->mol.nbf * ctx->mol.nbf;
    HF_HIP_CHECK(ctx, ctx->d_D.alloc(2 * nn));
    HF_HIP_CHECK(ctx, hipMemcpyAsync(ctx->d_D.p, D, ndens * nn * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    return hf_properties_impl(ctx, ndens, ctx->d_D.p, charges, dipole, ctx->stream);
}
