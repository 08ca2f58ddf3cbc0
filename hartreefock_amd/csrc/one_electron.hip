// One-electron matrices S, T, V on the device (the reference builds them on the host in
// HartreeFock/QuantumMatrix.cpp:34-155 from Obara-Saika recursions).  McMurchie-Davidson
// Hermite expansion, one thread per (shell pair, nucleus C), one launch per pair class (La, Lb)
// with La >= Lb so every array has a compile-time size and stays in registers; every primitive
// pair inside the thread; the nucleus-0 thread also writes S and T; V contributions are added with
// FP64 atomics into a zeroed V.  O(N^2 * atoms) work spread over that many threads.
//   S_ab = E^x_0 E^y_0 E^z_0 (pi/p)^{3/2}
//   T    = sum over directions of  b(2j+1) S_ij - 2b^2 S_i,j+2 - j(j-1)/2 S_i,j-2  (times the other two S)
//   V_ab = -sum_C Z_C 2pi/p sum_tuv E^x_t E^y_u E^z_v R_tuv(p, P - C)
// Writes full (symmetric) n x n row-major matrices.
#include "hf_internal.h"

namespace {

// F_m(T), m = 0..MMAX: table + Taylor + downward recursion, as the ERI kernels.
template <int MMAX>
__device__ __forceinline__ void boys_rt(double T, const double* __restrict__ tab, double* F) {
    using namespace hfgpu;
    if (T < BOYS_TMAX) {
        const int k = (int)rint(T * BOYS_INV_DT);
        const double x = k * BOYS_DT - T;
        const double* row = tab + k * BOYS_STRIDE + MMAX;
        double f = row[7];
#pragma unroll
        for (int j = 6; j >= 0; --j) f = row[j] + x * f * (1.0 / (j + 1));
        F[MMAX] = f;
        if (MMAX > 0) {
            const double e = exp(-T);
#pragma unroll
            for (int m = MMAX - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + e) * (1.0 / (2 * m + 1));
        }
    } else {
        F[0] = 0.88622692545275801365 / sqrt(T);
        if (MMAX > 0) {
            const double e = exp(-T);
#pragma unroll
            for (int m = 0; m < MMAX; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) * (0.5 / T);
        }
    }
}

// Hermite coefficients E[i][j][t] (i <= IM, j <= JM, t <= i + j) for one direction.
template <int IM, int JM>
__device__ __forceinline__ void hermite_e(double a, double b, double X, double (&E)[IM + 1][JM + 1][IM + JM + 1]) {
    const double p = a + b, q = a * b / p, o2p = 0.5 / p, qa = q * X / a, qb = q * X / b;
#pragma unroll
    for (int i = 0; i <= IM; ++i)
#pragma unroll
        for (int j = 0; j <= JM; ++j)
#pragma unroll
            for (int t = 0; t <= IM + JM; ++t) E[i][j][t] = 0.0;
    E[0][0][0] = exp(-q * X * X);
#pragma unroll
    for (int i = 0; i <= IM; ++i)
#pragma unroll
        for (int j = 0; j <= JM; ++j) {
            if (i == 0 && j == 0) continue;
#pragma unroll
            for (int t = 0; t <= i + j; ++t) {
                double v;
                if (j == 0) {
                    const int ii = i - 1;
                    v = -qa * E[ii][0][t] + (t + 1 <= ii ? (t + 1) * E[ii][0][t + 1] : 0.0);
                    if (t > 0) v += o2p * E[ii][0][t - 1];
                } else {
                    const int jj = j - 1;
                    v = qb * E[i][jj][t] + (t + 1 <= i + jj ? (t + 1) * E[i][jj][t + 1] : 0.0);
                    if (t > 0) v += o2p * E[i][jj][t - 1];
                }
                E[i][j][t] = v;
            }
        }
}

// Cartesian components of shells L = 0..2 in the reference order (x^l y^m z^n: xx, xy, xz, yy, yz, zz).
__device__ constexpr int kComp[3][6][3] = {
    {{0, 0, 0}},
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
    {{2, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2}}};

__host__ __device__ constexpr double dfact(int n) { return n > 1 ? n * dfact(n - 2) : 1.0; }

struct OneEArgs {
    int nshell, natoms, n;
    int atom0;               // nucleus of blockIdx.y = 0 in the launches after the first
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
    double* M;               // optional: dipole integrals about the origin, x, y, z (3 n x n)
};

__global__ void zero_kernel(double* __restrict__ x, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) x[i] = 0.0;
}

// FIRST: nucleus 0, which also computes S and T (and so needs E up to j = LB + 2); otherwise nucleus
// blockIdx.y + 1.
template <int LA, int LB, bool FIRST>
__global__ void __launch_bounds__(128) one_electron_kernel(OneEArgs a, const int2* __restrict__ pairs, int npairs) {
    constexpr int NA = (LA + 1) * (LA + 2) / 2, NB = (LB + 1) * (LB + 2) / 2, L = LA + LB, JM = FIRST ? LB + 2 : LB;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = FIRST ? 0 : blockIdx.y + a.atom0;     // nucleus
    if (t >= npairs) return;
    constexpr bool first = FIRST;
    const int sa = pairs[t].x, sb = pairs[t].y;
    const double* A = a.sxyz + 3 * sa;
    const double* B = a.sxyz + 3 * sb;
    const double AB[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
    const double C[3] = {a.axyz[3 * c], a.axyz[3 * c + 1], a.axyz[3 * c + 2]};
    double s[NA * NB], k[NA * NB], v[NA * NB];
    double mom[FIRST ? 3 : 1][NA * NB];
#pragma unroll
    for (int i = 0; i < NA * NB; ++i) {
        s[i] = k[i] = v[i] = 0.0;
#pragma unroll
        for (int d = 0; d < (FIRST ? 3 : 1); ++d) mom[d][i] = 0.0;
    }
    const bool want_m = FIRST && a.M != nullptr;
    const int pa0 = a.sprim0[sa], pb0 = a.sprim0[sb], npa = a.snprim[sa], npb = a.snprim[sb];
    for (int ia = 0; ia < npa; ++ia)
        for (int ib = 0; ib < npb; ++ib) {
            const double ea = a.alpha[pa0 + ia], eb = a.alpha[pb0 + ib], p = ea + eb;
            const double w = a.coef[pa0 + ia] * a.coef[pb0 + ib];
            double E[3][LA + 1][JM + 1][LA + JM + 1];
#pragma unroll
            for (int d = 0; d < 3; ++d) hermite_e<LA, JM>(ea, eb, AB[d], E[d]);
            const double sq = sqrt(M_PI / p);
            if constexpr (first) {
#pragma unroll
                for (int i = 0; i < NA; ++i)
#pragma unroll
                    for (int j = 0; j < NB; ++j) {
                        double s1[3], t1[3];
#pragma unroll
                        for (int d = 0; d < 3; ++d) {
                            const int li = kComp[LA][i][d], lj = kComp[LB][j][d];
                            const double s0 = E[d][li][lj][0] * sq;
                            const double sp2 = E[d][li][lj + 2 <= JM ? lj + 2 : JM][0] * sq;
                            const double sm2 = lj >= 2 ? E[d][li][lj >= 2 ? lj - 2 : 0][0] * sq : 0.0;
                            s1[d] = s0;
                            t1[d] = eb * (2 * lj + 1) * s0 - 2 * eb * eb * sp2 - 0.5 * lj * (lj - 1) * sm2;
                        }
                        s[i * NB + j] += w * s1[0] * s1[1] * s1[2];
                        k[i * NB + j] += w * (t1[0] * s1[1] * s1[2] + s1[0] * t1[1] * s1[2] + s1[0] * s1[1] * t1[2]);
                        if (want_m) {
                            // <a|x|b> about the origin: x = (x - P_x) + P_x, and the integral of
                            // (x - P_x) against the Hermite expansion keeps only t = 1:
                            // (E_1 + P_x E_0) sqrt(pi/p) in the moment direction.
#pragma unroll
                            for (int d = 0; d < (FIRST ? 3 : 1); ++d) {
                                const int li = kComp[LA][i][d], lj = kComp[LB][j][d];
                                const double Pd = (ea * A[d] + eb * B[d]) / p;
                                const double m1 = (E[d][li][lj][1] + Pd * E[d][li][lj][0]) * sq;
                                mom[d][i * NB + j] += w * m1 * (s1[(d + 1) % 3] * s1[(d + 2) % 3]);
                            }
                        }
                    }
            }
            const double PC[3] = {(ea * A[0] + eb * B[0]) / p - C[0], (ea * A[1] + eb * B[1]) / p - C[1],
                                  (ea * A[2] + eb * B[2]) / p - C[2]};
            double F[L + 1];
            boys_rt<L>(p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]), a.boys, F);
            // R[m][t][u][v] for t + u + v <= L - m (Hermite Coulomb integrals)
            double R[L + 1][L + 1][L + 1][L + 1];
            double fac = 1.0;
#pragma unroll
            for (int m = 0; m <= L; ++m) {
                R[m][0][0][0] = fac * F[m];
                fac *= -2.0 * p;
            }
#pragma unroll
            for (int m = L - 1; m >= 0; --m)
#pragma unroll
                for (int tot = 1; tot <= L - m; ++tot)
#pragma unroll
                    for (int tt = 0; tt <= tot; ++tt)
#pragma unroll
                        for (int uu = 0; uu <= tot - tt; ++uu) {
                            const int vv = tot - tt - uu;
                            double val;
                            if (tt > 0) {
                                val = PC[0] * R[m + 1][tt - 1][uu][vv];
                                if (tt > 1) val += (tt - 1) * R[m + 1][tt > 1 ? tt - 2 : 0][uu][vv];
                            } else if (uu > 0) {
                                val = PC[1] * R[m + 1][tt][uu - 1][vv];
                                if (uu > 1) val += (uu - 1) * R[m + 1][tt][uu > 1 ? uu - 2 : 0][vv];
                            } else {
                                val = PC[2] * R[m + 1][tt][uu][vv - 1];
                                if (vv > 1) val += (vv - 1) * R[m + 1][tt][uu][vv > 1 ? vv - 2 : 0];
                            }
                            R[m][tt][uu][vv] = val;
                        }
            const double zc = -a.Z[c] * w * 2.0 * M_PI / p;
#pragma unroll
            for (int i = 0; i < NA; ++i)
#pragma unroll
                for (int j = 0; j < NB; ++j) {
                    constexpr int X = 0, Y = 1, Z = 2;
                    const int lx = kComp[LA][i][X] + kComp[LB][j][X], ly = kComp[LA][i][Y] + kComp[LB][j][Y],
                              lz = kComp[LA][i][Z] + kComp[LB][j][Z];
                    double acc = 0;
#pragma unroll
                    for (int tt = 0; tt <= L; ++tt)
#pragma unroll
                        for (int uu = 0; uu <= L; ++uu)
#pragma unroll
                            for (int vv = 0; vv <= L; ++vv)
                                if (tt <= lx && uu <= ly && vv <= lz)
                                    acc += E[0][kComp[LA][i][X]][kComp[LB][j][X]][tt] *
                                           E[1][kComp[LA][i][Y]][kComp[LB][j][Y]][uu] *
                                           E[2][kComp[LA][i][Z]][kComp[LB][j][Z]][vv] * R[0][tt][uu][vv];
                    v[i * NB + j] += zc * acc;
                }
        }
    const int n = a.n, a0 = a.sbf[sa], b0 = a.sbf[sb];
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const double f = 1.0 / sqrt(dfact(2 * kComp[LA][i][0] - 1) * dfact(2 * kComp[LA][i][1] - 1) *
                                        dfact(2 * kComp[LA][i][2] - 1) * dfact(2 * kComp[LB][j][0] - 1) *
                                        dfact(2 * kComp[LB][j][1] - 1) * dfact(2 * kComp[LB][j][2] - 1));
            const size_t r = (size_t)(a0 + i), cc = (size_t)(b0 + j);
            if constexpr (first) {
                a.S[r * n + cc] = a.S[cc * n + r] = f * s[i * NB + j];
                a.T[r * n + cc] = a.T[cc * n + r] = f * k[i * NB + j];
                if (want_m) {
                    const size_t nn = (size_t)n * n;
#pragma unroll
                    for (int d = 0; d < (FIRST ? 3 : 1); ++d)
                        a.M[d * nn + r * n + cc] = a.M[d * nn + cc * n + r] = f * mom[d][i * NB + j];
                }
            }
            unsafeAtomicAdd(&a.V[r * n + cc], f * v[i * NB + j]);
            if (sa != sb) unsafeAtomicAdd(&a.V[cc * n + r], f * v[i * NB + j]);
        }
}

}  // namespace

// Device S, T, V for the context's molecule (row-major n x n device buffers), on stream st.
int hf_one_electron_device(hf_ctx* ctx, double* dS, double* dT, double* dV, hipStream_t st, double* dM) {
    const hfhost::Molecule& mol = ctx->mol;
    const int ns = (int)mol.shells.size();
    const bool rebuild = !ctx->oe_valid;   // inputs of the current molecule not uploaded yet
    bool uploaded = false;
    if (rebuild) {
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
        std::vector<int> ints;
        for (const std::vector<int>* v : {&L, &bf, &p0, &np}) ints.insert(ints.end(), v->begin(), v->end());
        std::vector<double> dbl;
        for (const std::vector<double>* v : {(const std::vector<double>*)&al, (const std::vector<double>*)&co,
                                             (const std::vector<double>*)&Z, &mol.xyz})
            dbl.insert(dbl.end(), v->begin(), v->end());
        if (int rc = hf_upload(ctx, ip, ints.data(), ints.size() * sizeof(int))) return rc;
        if (int rc = hf_upload(ctx, ctx->d_oe_prim.p, dbl.data(), dbl.size() * sizeof(double))) return rc;
        uploaded = true;
        ctx->oe_nprim = al.size();
    }
    OneEArgs a;
    a.atom0 = 1;
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
    a.M = dM;
    // shell pairs by class (La, Lb), La >= Lb, built once per molecule
    constexpr int kOeClasses = 6;
    static const int kOeL[kOeClasses][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}};
    if (rebuild) {
        std::vector<std::vector<int2>> byc(kOeClasses);
        for (int x = 0; x < ns; ++x)
            for (int y = 0; y <= x; ++y) {
                int sa = x, sb = y;
                if (mol.shells[sa].L < mol.shells[sb].L) std::swap(sa, sb);
                const int La = mol.shells[sa].L, Lb = mol.shells[sb].L;
                for (int c = 0; c < kOeClasses; ++c)
                    if (kOeL[c][0] == La && kOeL[c][1] == Lb) byc[c].push_back(make_int2(sa, sb));
            }
        std::vector<int2> all;
        ctx->oe_class_off.assign(kOeClasses + 1, 0);
        for (int c = 0; c < kOeClasses; ++c) {
            ctx->oe_class_off[c] = (int)all.size();
            all.insert(all.end(), byc[c].begin(), byc[c].end());
        }
        ctx->oe_class_off[kOeClasses] = (int)all.size();
        HF_HIP_CHECK(ctx, ctx->d_oe_pairs.alloc(std::max<size_t>(all.size(), 1)));
        if (!all.empty()) {
            if (int rc = hf_upload(ctx, ctx->d_oe_pairs.p, all.data(), all.size() * sizeof(int2))) return rc;
            uploaded = true;
        }
        ctx->oe_valid = true;
    }
    // the uploads are queued on the context's stream
    if (uploaded && st != ctx->stream) HF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    const size_t nn = (size_t)mol.nbf * mol.nbf;
    // Deterministic mode: one launch per nucleus, in order, so that every V element receives its nuclei's
    // terms in the same order each time (within one launch each shell pair's elements have one writer).
    const bool det = ctx->opt.deterministic != 0;
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)std::min<size_t>((nn + 255) / 256, 4096)), dim3(256), 0, st, dV, nn);
    for (int c = 0; c < kOeClasses && a.natoms > 0; ++c) {
        const int off = ctx->oe_class_off[c], np = ctx->oe_class_off[c + 1] - off;
        if (np == 0) continue;
        const int2* pl = ctx->d_oe_pairs.p + off;
        const dim3 g1((unsigned)((np + 127) / 128), 1), gr((unsigned)((np + 127) / 128), (unsigned)(a.natoms - 1));
        switch (c) {
#define HF_OE_LAUNCH(C, LA, LB)                                                                              \
    case C:                                                                                                  \
        hipLaunchKernelGGL((one_electron_kernel<LA, LB, true>), g1, dim3(128), 0, st, a, pl, np);           \
        if (a.natoms > 1 && !det) {                                                                          \
            a.atom0 = 1;                                                                                     \
            hipLaunchKernelGGL((one_electron_kernel<LA, LB, false>), gr, dim3(128), 0, st, a, pl, np);       \
        }                                                                                                    \
        for (int at = 1; det && at < a.natoms; ++at) {                                                       \
            a.atom0 = at;                                                                                    \
            hipLaunchKernelGGL((one_electron_kernel<LA, LB, false>), g1, dim3(128), 0, st, a, pl, np);       \
        }                                                                                                    \
        break;
            HF_OE_LAUNCH(0, 0, 0)
            HF_OE_LAUNCH(1, 1, 0)
            HF_OE_LAUNCH(2, 1, 1)
            HF_OE_LAUNCH(3, 2, 0)
            HF_OE_LAUNCH(4, 2, 1)
            HF_OE_LAUNCH(5, 2, 2)
#undef HF_OE_LAUNCH
        }
    }
    HF_HIP_CHECK(ctx, hipGetLastError());
    return HF_OK;
}
