// One-electron matrices S, T, V on the device (the reference builds them on the host in
// HartreeFock/QuantumMatrix.cpp:34-155 from Obara-Saika recursions).  McMurchie-Davidson
// Hermite expansion, one thread per (shell pair a >= b, nucleus C): every primitive pair inside
// the thread; the nucleus-0 thread also writes S and T; V contributions are added with FP64
// atomics into a zeroed V.  O(N^2 * atoms) work spread over that many threads.
//   S_ab = E^x_0 E^y_0 E^z_0 (pi/p)^{3/2}
//   T    = sum over directions of  b(2j+1) S_ij - 2b^2 S_i,j+2 - j(j-1)/2 S_i,j-2  (times the other two S)
//   V_ab = -sum_C Z_C 2pi/p sum_tuv E^x_t E^y_u E^z_v R_tuv(p, P - C)
// Writes full (symmetric) n x n row-major matrices.
#include "hf_internal.h"

namespace {

constexpr int kI = 3;   // i <= 2
constexpr int kJ = 5;   // j <= 2 + 2
constexpr int kT = 7;   // t <= i + j
constexpr int kL = 5;   // t+u+v <= 4

// F_m(T), m = 0..mmax (mmax <= 4): table + Taylor + downward recursion, as the ERI kernels.
__device__ void boys_rt(double T, int mmax, const double* __restrict__ tab, double* F) {
    using namespace hfgpu;
    if (T < BOYS_TMAX) {
        const int k = (int)rint(T * BOYS_INV_DT);
        const double x = k * BOYS_DT - T;
        const double* row = tab + k * BOYS_STRIDE + mmax;
        double f = row[7];
        for (int j = 6; j >= 0; --j) f = row[j] + x * f / (j + 1);
        const double e = exp(-T);
        F[mmax] = f;
        for (int m = mmax - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + e) / (2 * m + 1);
    } else {
        const double e = exp(-T);
        F[0] = 0.88622692545275801365 / sqrt(T);
        for (int m = 0; m < mmax; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) * (0.5 / T);
    }
}

__device__ void hermite_e(double a, double b, double X, int imax, int jmax, double (&E)[kI][kJ][kT]) {
    const double p = a + b, q = a * b / p, o2p = 0.5 / p;
    for (int i = 0; i < kI; ++i)
        for (int j = 0; j < kJ; ++j)
            for (int t = 0; t < kT; ++t) E[i][j][t] = 0.0;
    E[0][0][0] = exp(-q * X * X);
    for (int i = 0; i <= imax; ++i)
        for (int j = 0; j <= jmax; ++j) {
            if (i == 0 && j == 0) continue;
            for (int t = 0; t <= i + j; ++t) {
                double v;
                if (j == 0) {
                    const int ii = i - 1;
                    v = -(q * X / a) * E[ii][0][t] + (t + 1 <= ii ? (t + 1) * E[ii][0][t + 1] : 0.0);
                    if (t > 0) v += o2p * E[ii][0][t - 1];
                } else {
                    const int jj = j - 1;
                    v = (q * X / b) * E[i][jj][t] + (t + 1 <= i + jj ? (t + 1) * E[i][jj][t + 1] : 0.0);
                    if (t > 0) v += o2p * E[i][jj][t - 1];
                }
                E[i][j][t] = v;
            }
        }
}

__device__ __forceinline__ void comp_of(int L, int k, int& l, int& m, int& n) {
    // k-th Cartesian component of shell L in the order xx,xy,xz,yy,yz,zz
    int idx = 0;
    for (int a = L; a >= 0; --a)
        for (int b = L - a; b >= 0; --b) {
            if (idx == k) { l = a; m = b; n = L - a - b; return; }
            ++idx;
        }
    l = m = n = 0;
}

__device__ __forceinline__ double dfact(int n) {
    double r = 1;
    for (; n > 1; n -= 2) r *= n;
    return r;
}

struct OneEArgs {
    int nshell, natoms, n;
    const double* sxyz;
    const int* sL;
    const int* sbf;
    const int* sprim0;
    const int* snprim;
    const double* alpha;
    const double* coef;
    const double* axyz;
    const double* Z;
    const double* boys;
    double *S, *T, *V;
};

__global__ void zero_kernel(double* __restrict__ x, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) x[i] = 0.0;
}

__global__ void __launch_bounds__(128) one_electron_kernel(OneEArgs a) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t npairs = (int64_t)a.nshell * (a.nshell + 1) / 2;
    const int c = blockIdx.y;           // nucleus
    if (t >= npairs) return;
    const bool first = c == 0;
    // t -> (sa >= sb)
    int sa = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((int64_t)sa * (sa + 1) / 2 > t) --sa;
    while ((int64_t)(sa + 1) * (sa + 2) / 2 <= t) ++sa;
    const int sb = (int)(t - (int64_t)sa * (sa + 1) / 2);
    const int La = a.sL[sa], Lb = a.sL[sb];
    const int na = (La + 1) * (La + 2) / 2, nb = (Lb + 1) * (Lb + 2) / 2;
    const double* A = a.sxyz + 3 * sa;
    const double* B = a.sxyz + 3 * sb;
    const double AB[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
    double s[36], k[36], v[36];
    for (int i = 0; i < 36; ++i) s[i] = k[i] = v[i] = 0.0;
    const int L = La + Lb;
    for (int ia = 0; ia < a.snprim[sa]; ++ia)
        for (int ib = 0; ib < a.snprim[sb]; ++ib) {
            const double ea = a.alpha[a.sprim0[sa] + ia], eb = a.alpha[a.sprim0[sb] + ib], p = ea + eb;
            const double w = a.coef[a.sprim0[sa] + ia] * a.coef[a.sprim0[sb] + ib];
            double E[3][kI][kJ][kT];
            for (int d = 0; d < 3; ++d) hermite_e(ea, eb, AB[d], La, first ? Lb + 2 : Lb, E[d]);
            const double sq = sqrt(M_PI / p);
            const double P[3] = {(ea * A[0] + eb * B[0]) / p, (ea * A[1] + eb * B[1]) / p, (ea * A[2] + eb * B[2]) / p};
            if (first) {
                for (int i = 0; i < na; ++i)
                    for (int j = 0; j < nb; ++j) {
                        int li[3], lj[3];
                        comp_of(La, i, li[0], li[1], li[2]);
                        comp_of(Lb, j, lj[0], lj[1], lj[2]);
                        double s1[3], t1[3];
                        for (int d = 0; d < 3; ++d) {
                            const double s0 = E[d][li[d]][lj[d]][0] * sq;
                            const double sp2 = E[d][li[d]][lj[d] + 2][0] * sq;
                            const double sm2 = lj[d] >= 2 ? E[d][li[d]][lj[d] - 2][0] * sq : 0.0;
                            s1[d] = s0;
                            t1[d] = eb * (2 * lj[d] + 1) * s0 - 2 * eb * eb * sp2 - 0.5 * lj[d] * (lj[d] - 1) * sm2;
                        }
                        s[i * nb + j] += w * s1[0] * s1[1] * s1[2];
                        k[i * nb + j] += w * (t1[0] * s1[1] * s1[2] + s1[0] * t1[1] * s1[2] + s1[0] * s1[1] * t1[2]);
                    }
            }
            {
                const double PC[3] = {P[0] - a.axyz[3 * c], P[1] - a.axyz[3 * c + 1], P[2] - a.axyz[3 * c + 2]};
                double F[kL];
                boys_rt(p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]), L, a.boys, F);
                double R[kL][kL][kL][kL];
                double fac = 1.0;
                for (int m = 0; m <= L; ++m) {
                    R[m][0][0][0] = fac * F[m];
                    fac *= -2.0 * p;
                }
                for (int m = L - 1; m >= 0; --m)
                    for (int tot = 1; tot <= L - m; ++tot)
                        for (int tt = 0; tt <= tot; ++tt)
                            for (int uu = 0; uu <= tot - tt; ++uu) {
                                const int vv = tot - tt - uu;
                                double val;
                                if (tt > 0) {
                                    val = PC[0] * R[m + 1][tt - 1][uu][vv];
                                    if (tt > 1) val += (tt - 1) * R[m + 1][tt - 2][uu][vv];
                                } else if (uu > 0) {
                                    val = PC[1] * R[m + 1][tt][uu - 1][vv];
                                    if (uu > 1) val += (uu - 1) * R[m + 1][tt][uu - 2][vv];
                                } else {
                                    val = PC[2] * R[m + 1][tt][uu][vv - 1];
                                    if (vv > 1) val += (vv - 1) * R[m + 1][tt][uu][vv - 2];
                                }
                                R[m][tt][uu][vv] = val;
                            }
                const double zc = -a.Z[c] * w * 2.0 * M_PI / p;
                for (int i = 0; i < na; ++i)
                    for (int j = 0; j < nb; ++j) {
                        int li[3], lj[3];
                        comp_of(La, i, li[0], li[1], li[2]);
                        comp_of(Lb, j, lj[0], lj[1], lj[2]);
                        double acc = 0;
                        for (int tt = 0; tt <= li[0] + lj[0]; ++tt)
                            for (int uu = 0; uu <= li[1] + lj[1]; ++uu)
                                for (int vv = 0; vv <= li[2] + lj[2]; ++vv)
                                    acc += E[0][li[0]][lj[0]][tt] * E[1][li[1]][lj[1]][uu] * E[2][li[2]][lj[2]][vv] *
                                           R[0][tt][uu][vv];
                        v[i * nb + j] += zc * acc;
                    }
            }
        }
    const int n = a.n, a0 = a.sbf[sa], b0 = a.sbf[sb];
    for (int i = 0; i < na; ++i)
        for (int j = 0; j < nb; ++j) {
            int li[3], lj[3];
            comp_of(La, i, li[0], li[1], li[2]);
            comp_of(Lb, j, lj[0], lj[1], lj[2]);
            const double f = 1.0 / sqrt(dfact(2 * li[0] - 1) * dfact(2 * li[1] - 1) * dfact(2 * li[2] - 1) *
                                        dfact(2 * lj[0] - 1) * dfact(2 * lj[1] - 1) * dfact(2 * lj[2] - 1));
            const size_t r = (size_t)(a0 + i), cc = (size_t)(b0 + j);
            if (first) {
                a.S[r * n + cc] = a.S[cc * n + r] = f * s[i * nb + j];
                a.T[r * n + cc] = a.T[cc * n + r] = f * k[i * nb + j];
            }
            unsafeAtomicAdd(&a.V[r * n + cc], f * v[i * nb + j]);
            if (sa != sb) unsafeAtomicAdd(&a.V[cc * n + r], f * v[i * nb + j]);
        }
}

}  // namespace

// Device S, T, V for the context's molecule (row-major n x n device buffers), on stream st.
int hf_one_electron_device(hf_ctx* ctx, double* dS, double* dT, double* dV, hipStream_t st) {
    const hfhost::Molecule& mol = ctx->mol;
    const int ns = (int)mol.shells.size();
    if (!ctx->d_oe_L.p) {
        std::vector<int> L(ns), bf(ns), p0(ns), np(ns);
        std::vector<double> al, co;
        for (int s = 0; s < ns; ++s) {
            L[s] = mol.shells[s].L;
            bf[s] = mol.shells[s].bf;
            p0[s] = (int)al.size();
            np[s] = (int)mol.shells[s].alpha.size();
            al.insert(al.end(), mol.shells[s].alpha.begin(), mol.shells[s].alpha.end());
            co.insert(co.end(), mol.shells[s].coef.begin(), mol.shells[s].coef.end());
        }
        std::vector<double> Z(mol.Z.begin(), mol.Z.end());
        HF_HIP_CHECK(ctx, ctx->d_oe_L.alloc(4 * (size_t)ns));
        HF_HIP_CHECK(ctx, ctx->d_oe_prim.alloc(2 * al.size() + Z.size() + mol.xyz.size()));
        int* ip = ctx->d_oe_L.p;
        HF_HIP_CHECK(ctx, hipMemcpy(ip, L.data(), ns * sizeof(int), hipMemcpyHostToDevice));
        HF_HIP_CHECK(ctx, hipMemcpy(ip + ns, bf.data(), ns * sizeof(int), hipMemcpyHostToDevice));
        HF_HIP_CHECK(ctx, hipMemcpy(ip + 2 * ns, p0.data(), ns * sizeof(int), hipMemcpyHostToDevice));
        HF_HIP_CHECK(ctx, hipMemcpy(ip + 3 * ns, np.data(), ns * sizeof(int), hipMemcpyHostToDevice));
        double* dp = ctx->d_oe_prim.p;
        HF_HIP_CHECK(ctx, hipMemcpy(dp, al.data(), al.size() * sizeof(double), hipMemcpyHostToDevice));
        HF_HIP_CHECK(ctx, hipMemcpy(dp + al.size(), co.data(), co.size() * sizeof(double), hipMemcpyHostToDevice));
        HF_HIP_CHECK(ctx, hipMemcpy(dp + 2 * al.size(), Z.data(), Z.size() * sizeof(double), hipMemcpyHostToDevice));
        HF_HIP_CHECK(ctx, hipMemcpy(dp + 2 * al.size() + Z.size(), mol.xyz.data(), mol.xyz.size() * sizeof(double),
                                    hipMemcpyHostToDevice));
        ctx->oe_nprim = al.size();
    }
    OneEArgs a;
    a.nshell = ns;
    a.natoms = (int)mol.Z.size();
    a.n = mol.nbf;
    a.sxyz = ctx->d_shell_xyz.p;
    a.sL = ctx->d_oe_L.p;
    a.sbf = ctx->d_oe_L.p + ns;
    a.sprim0 = ctx->d_oe_L.p + 2 * ns;
    a.snprim = ctx->d_oe_L.p + 3 * ns;
    a.alpha = ctx->d_oe_prim.p;
    a.coef = ctx->d_oe_prim.p + ctx->oe_nprim;
    a.Z = ctx->d_oe_prim.p + 2 * ctx->oe_nprim;
    a.axyz = a.Z + mol.Z.size();
    a.boys = ctx->d_boys.p;
    a.S = dS;
    a.T = dT;
    a.V = dV;
    const int64_t npairs = (int64_t)ns * (ns + 1) / 2;
    const size_t nn = (size_t)mol.nbf * mol.nbf;
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)std::min<size_t>((nn + 255) / 256, 4096)), dim3(256), 0, st, dV, nn);
    if (a.natoms > 0)
        hipLaunchKernelGGL(one_electron_kernel, dim3((unsigned)((npairs + 127) / 128), (unsigned)a.natoms), dim3(128), 0,
                           st, a);
    HF_HIP_CHECK(ctx, hipGetLastError());
    return HF_OK;
}
