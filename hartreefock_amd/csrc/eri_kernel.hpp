// Shell-quartet ERI kernel, one template instance per angular class and sink.
//
// One thread evaluates one contracted shell quartet (ab|cd):
//   for every primitive pair of the bra x every primitive pair of the ket:
//     Boys F_0..F_L (table + Taylor, FP64), HGP vertical recursion and electron
//     transfer (generated straight-line code), accumulated into the contracted
//     [e0|f0] block held in registers;
//   then the horizontal recursions on bra and ket (generated) give the Cartesian
//   block (ab|cd), which goes to one of three sinks:
//     MODE_JK      digestion into J and K (8-fold symmetry): exchange rows of the block's bra
//                  pairs staged in LDS, Coulomb ket block with FP64 global atomics,
//     MODE_PACKED  the reference's packed unique-ERI store,
//     MODE_DIAG    the Schwarz bound sqrt(max |(ab|ab)|) of a shell pair.
#pragma once

#include <algorithm>

#include "hf_device.h"
#include "wave_reduce.h"
#include "generated/eri_classes.inc"

namespace hfgpu {

constexpr double kTwoPi25 = 34.98683665524972497;   // 2 pi^{5/2}
constexpr double kSqrtPiHalf = 0.88622692545275801365;  // sqrt(pi)/2
// Classes with at most this many VRR accumulators evaluate bra primitives two at a time.
constexpr int kKetPairMaxAcc = 27;
// J/K kernels of classes with L up to kBoysLdsMaxL and a bra pair of at most kBoysLdsMaxBra basis
// functions stage their slice of the Boys table (F_L..F_{L+7} and e^{-T_k}, kBoysLdsStride doubles
// per row, 26 KB) in LDS; larger bras need the LDS for their exchange-row slots.
constexpr int kBoysLdsMaxL = 3;
constexpr int kBoysLdsMaxBra = 6;
template <class C>
constexpr bool boys_in_lds() { return C::L <= kBoysLdsMaxL && C::NA + C::NB <= kBoysLdsMaxBra && !C::TRUNCATED; }
constexpr int kBoysLdsStride = 9;
constexpr size_t kBoysLdsBytes = (size_t)BOYS_NROW * kBoysLdsStride * sizeof(double);
constexpr int kBoysLdsStride32 = 7;    // FP32 tier: F_L..F_{L+5}

// 1/sqrt(x) for normal positive x: the hardware estimate refined by two Newton steps
// r <- r (3 - x r^2) / 2 (each step squares the relative error), on the device; exact on the host.
__host__ __device__ __forceinline__ double rsqrt_f64(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rsq(x);
    r = r * (1.5 - 0.5 * x * r * r);
    r = r * (1.5 - 0.5 * x * r * r);
    return r;
#else
    return 1.0 / sqrt(x);
#endif
}

__host__ __device__ __forceinline__ float rsqrt_f32(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / sqrtf(x);
#endif
}

// F_m(T) * pref for m = 0..L.  Below BOYS_TMAX: F_L by an 8-term Taylor expansion about the nearest
// grid point T_k (|x| = |T_k - T| <= 0.05, truncation below 1e-17 relative) and downward
// recursion F_m = (2T F_{m+1} + e^{-T}) / (2m+1), with e^{-T} = e^{-T_k} e^x from the table and
// a 9-term series for e^x.  Beyond: F_0 = sqrt(pi/T)/2 and upward recursion.
// STRIDE, C0, ECOL: row stride of the table, column of F_L, column of e^{-T_k} (the device table, or
// the class's slice of it staged in LDS).
template <int L, int STRIDE = BOYS_STRIDE, int C0 = L, int ECOL = BOYS_NCOL>
__host__ __device__ __forceinline__ void boys_scaled(double T, double pref, const double* __restrict__ tab, double* F) {
    if (T < BOYS_TMAX) {
        const int k = (int)rint(T * BOYS_INV_DT);
        const double x = k * BOYS_DT - T;          // F_m(T) = sum_j F_{m+j}(T_k) x^j / j!
        const double* row = tab + k * STRIDE;
        double f = row[C0 + 7];
        f = row[C0 + 6] + x * f * (1.0 / 7.0);
        f = row[C0 + 5] + x * f * (1.0 / 6.0);
        f = row[C0 + 4] + x * f * (1.0 / 5.0);
        f = row[C0 + 3] + x * f * (1.0 / 4.0);
        f = row[C0 + 2] + x * f * (1.0 / 3.0);
        f = row[C0 + 1] + x * f * (1.0 / 2.0);
        f = row[C0] + x * f;
        F[L] = f;
        if (L > 0) {
            double ex = 1.0 + x * (1.0 / 8.0);
            ex = 1.0 + x * (1.0 / 7.0) * ex;
            ex = 1.0 + x * (1.0 / 6.0) * ex;
            ex = 1.0 + x * (1.0 / 5.0) * ex;
            ex = 1.0 + x * (1.0 / 4.0) * ex;
            ex = 1.0 + x * (1.0 / 3.0) * ex;
            ex = 1.0 + x * (1.0 / 2.0) * ex;
            ex = 1.0 + x * ex;
            const double e = row[ECOL] * ex;
#pragma unroll
            for (int m = L - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + e) * (1.0 / (2 * m + 1));
        }
    } else {
        const double r = rsqrt_f64(T), inv2T = 0.5 * r * r;
        F[0] = kSqrtPiHalf * r;
        if (L > 0) {
            // e^{-T} < 2.4e-16 here and enters F_{m+1} = ((2m+1) F_m - e^{-T}) / 2T only as a correction
            // below 1e-11 of (2m+1) F_m for m <= 8, so FP32 exp (relative error ~1e-7) is exact to
            // FP64 rounding and avoids the FP64 exp polynomial
#if defined(__HIP_DEVICE_COMPILE__)
            const double e = (double)__expf(-(float)T);
#else
            const double e = exp(-T);
#endif
#pragma unroll
            for (int m = 0; m < L; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) * inv2T;
        }
    }
#pragma unroll
    for (int m = 0; m <= L; ++m) F[m] *= pref;
}

// FP32 evaluation of F_m(T) * pref (quartets of the mixed-precision tier): a 6-term Taylor
// expansion of F_L from an FP32 copy of the table (truncation < 3e-9 relative for |x| <= 0.05),
// FP32 exp and downward recursion below BOYS_TMAX; above it the asymptotic form
// F_0 = sqrt(pi/T)/2 and the upward recursion F_{m+1} = (2m+1) F_m / (2T) in FP64, rounded.  The
// e^{-T} term of that recursion is below 1.6e-15 of (2m+1) F_m for T >= 36 and is left out.
template <int L, int STRIDE = BOYS_STRIDE, int C0 = L>
__host__ __device__ __forceinline__ void boys_scaled_f32(float T, float pref, const float* __restrict__ tab, float* F) {
    if (T < (float)BOYS_TMAX) {
        const int k = (int)rintf(T * (float)BOYS_INV_DT);
        const float x = (float)k * (float)BOYS_DT - T;
        const float* row = tab + k * STRIDE + C0;
        float f = row[5];
        f = row[4] + x * f * (1.0f / 5.0f);
        f = row[3] + x * f * (1.0f / 4.0f);
        f = row[2] + x * f * (1.0f / 3.0f);
        f = row[1] + x * f * (1.0f / 2.0f);
        f = row[0] + x * f;
        F[L] = f;
        if (L > 0) {
            const float e = expf(-T);
#pragma unroll
            for (int m = L - 1; m >= 0; --m) F[m] = (2.0f * T * F[m + 1] + e) * (1.0f / (2 * m + 1));
        }
#pragma unroll
        for (int m = 0; m <= L; ++m) F[m] *= pref;
    } else {
        const double Td = T, r = rsqrt_f64(Td), inv2T = 0.5 * r * r;
        double Fd = kSqrtPiHalf * r;
        F[0] = (float)(Fd * pref);
#pragma unroll
        for (int m = 0; m < L; ++m) {
            Fd = (2 * m + 1) * Fd * inv2T;
            F[m + 1] = (float)(Fd * pref);
        }
    }
}

__host__ __device__ __forceinline__ double rsqrt_r(double x) { return rsqrt_f64(x); }
__host__ __device__ __forceinline__ float rsqrt_r(float x) { return rsqrt_f32(x); }

__device__ __forceinline__ uint64_t two_index(uint64_t i, uint64_t j) {
    return i < j ? j * (j + 1) / 2 + i : i * (i + 1) / 2 + j;
}

template <class C, typename V>
__device__ __forceinline__ void sink_jk(const V* __restrict__ v, int a0, int b0, int c0, int d0, double w,
                                        const SinkArgs& s) {
    constexpr int NA = C::NA, NB = C::NB, NC = C::NC, ND = C::ND;
    const int n = s.n;
    const size_t nn = (size_t)n * n;
    // Coulomb: J'_ab += 2w sum_cd (ab|cd) D_cd ;  J'_cd += 2w sum_ab (ab|cd) D_ab  (D summed over densities)
    {
        double Dcd[NC * ND], Dab[NA * NB];
#pragma unroll
        for (int i = 0; i < NC; ++i)
#pragma unroll
            for (int j = 0; j < ND; ++j) {
                double x = 0;
                for (int k = 0; k < s.ndens; ++k) x += s.D[k * nn + (size_t)(c0 + i) * n + d0 + j];
                Dcd[i * ND + j] = x;
            }
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                double x = 0;
                for (int k = 0; k < s.ndens; ++k) x += s.D[k * nn + (size_t)(a0 + i) * n + b0 + j];
                Dab[i * NB + j] = x;
            }
        const double w2 = 2.0 * w;
#pragma unroll
        for (int ia = 0; ia < NA; ++ia)
#pragma unroll
            for (int ib = 0; ib < NB; ++ib) {
                double acc = 0;
#pragma unroll
                for (int ic = 0; ic < NC; ++ic)
#pragma unroll
                    for (int id = 0; id < ND; ++id) acc += v[((ia * NB + ib) * NC + ic) * ND + id] * Dcd[ic * ND + id];
                sink_add_J(s, (size_t)(a0 + ia) * n + b0 + ib, w2 * acc);
            }
#pragma unroll
        for (int ic = 0; ic < NC; ++ic)
#pragma unroll
            for (int id = 0; id < ND; ++id) {
                double acc = 0;
#pragma unroll
                for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                    for (int ib = 0; ib < NB; ++ib) acc += v[((ia * NB + ib) * NC + ic) * ND + id] * Dab[ia * NB + ib];
                sink_add_J(s, (size_t)(c0 + ic) * n + d0 + id, w2 * acc);
            }
    }
    // Exchange, per density: K'_ac += w (ab|cd) D_bd, K'_ad += w D_bc, K'_bc += w D_ad, K'_bd += w D_ac
    for (int k = 0; k < s.ndens; ++k) {
        const double* __restrict__ D = s.D + k * nn;
        double Dbd[NB * ND], Dbc[NB * NC], Dad[NA * ND], Dac[NA * NC];
#pragma unroll
        for (int i = 0; i < NB; ++i) {
#pragma unroll
            for (int j = 0; j < ND; ++j) Dbd[i * ND + j] = D[(size_t)(b0 + i) * n + d0 + j];
#pragma unroll
            for (int j = 0; j < NC; ++j) Dbc[i * NC + j] = D[(size_t)(b0 + i) * n + c0 + j];
        }
#pragma unroll
        for (int i = 0; i < NA; ++i) {
#pragma unroll
            for (int j = 0; j < ND; ++j) Dad[i * ND + j] = D[(size_t)(a0 + i) * n + d0 + j];
#pragma unroll
            for (int j = 0; j < NC; ++j) Dac[i * NC + j] = D[(size_t)(a0 + i) * n + c0 + j];
        }
#pragma unroll
        for (int ia = 0; ia < NA; ++ia)
#pragma unroll
            for (int ic = 0; ic < NC; ++ic) {
                double acc = 0;
#pragma unroll
                for (int ib = 0; ib < NB; ++ib)
#pragma unroll
                    for (int id = 0; id < ND; ++id) acc += v[((ia * NB + ib) * NC + ic) * ND + id] * Dbd[ib * ND + id];
                sink_add_K(s, k, (size_t)(a0 + ia) * n + c0 + ic, w * acc);
            }
#pragma unroll
        for (int ia = 0; ia < NA; ++ia)
#pragma unroll
            for (int id = 0; id < ND; ++id) {
                double acc = 0;
#pragma unroll
                for (int ib = 0; ib < NB; ++ib)
#pragma unroll
                    for (int ic = 0; ic < NC; ++ic) acc += v[((ia * NB + ib) * NC + ic) * ND + id] * Dbc[ib * NC + ic];
                sink_add_K(s, k, (size_t)(a0 + ia) * n + d0 + id, w * acc);
            }
#pragma unroll
        for (int ib = 0; ib < NB; ++ib)
#pragma unroll
            for (int ic = 0; ic < NC; ++ic) {
                double acc = 0;
#pragma unroll
                for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                    for (int id = 0; id < ND; ++id) acc += v[((ia * NB + ib) * NC + ic) * ND + id] * Dad[ia * ND + id];
                sink_add_K(s, k, (size_t)(b0 + ib) * n + c0 + ic, w * acc);
            }
#pragma unroll
        for (int ib = 0; ib < NB; ++ib)
#pragma unroll
            for (int id = 0; id < ND; ++id) {
                double acc = 0;
#pragma unroll
                for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                    for (int ic = 0; ic < NC; ++ic) acc += v[((ia * NB + ib) * NC + ic) * ND + id] * Dac[ia * NC + ic];
                sink_add_K(s, k, (size_t)(b0 + ib) * n + d0 + id, w * acc);
            }
    }
}

// Digestion of one quartet whose bra pair has an LDS slot.  The slot holds, per density k, the
// exchange rows of the bra's functions: NA rows a0.. then NB rows b0.., n columns each; the four
// exchange blocks K'_ac, K'_ad, K'_bc, K'_bd all land in those rows.  The ket-side Coulomb block
// J'_cd goes to global memory; the bra-side Coulomb block J'_ab is returned in jab for a
// reduction across the wavefront.
template <class C, typename V>
__device__ __forceinline__ void sink_jk_rows(const V* __restrict__ v, double* __restrict__ rows, int a0, int b0,
                                             int c0, int d0, double w, const SinkArgs& s, double* __restrict__ jab) {
    constexpr int NA = C::NA, NB = C::NB, NC = C::NC, ND = C::ND;
    const int n = s.n;
    const size_t nn = (size_t)n * n;
    {
        double Dcd[NC * ND], Dab[NA * NB];
#pragma unroll
        for (int i = 0; i < NC; ++i)
#pragma unroll
            for (int j = 0; j < ND; ++j) {
                double x = 0;
                for (int k = 0; k < s.ndens; ++k) x += s.D[k * nn + (size_t)(c0 + i) * n + d0 + j];
                Dcd[i * ND + j] = x;
            }
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                double x = 0;
                for (int k = 0; k < s.ndens; ++k) x += s.D[k * nn + (size_t)(a0 + i) * n + b0 + j];
                Dab[i * NB + j] = x;
            }
        const double w2 = 2.0 * w;
#pragma unroll
        for (int ia = 0; ia < NA; ++ia)
#pragma unroll
            for (int ib = 0; ib < NB; ++ib) {
                double acc = 0;
#pragma unroll
                for (int ic = 0; ic < NC; ++ic)
#pragma unroll
                    for (int id = 0; id < ND; ++id) acc += v[((ia * NB + ib) * NC + ic) * ND + id] * Dcd[ic * ND + id];
                jab[ia * NB + ib] = w2 * acc;
            }
#pragma unroll
        for (int ic = 0; ic < NC; ++ic)
#pragma unroll
            for (int id = 0; id < ND; ++id) {
                double acc = 0;
#pragma unroll
                for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                    for (int ib = 0; ib < NB; ++ib) acc += v[((ia * NB + ib) * NC + ic) * ND + id] * Dab[ia * NB + ib];
                sink_add_J(s, (size_t)(c0 + ic) * n + d0 + id, w2 * acc);
            }
    }
    for (int k = 0; k < s.ndens; ++k) {
        const double* __restrict__ D = s.D + k * nn;
        double* ra = rows + (size_t)k * (NA + NB) * n;
        double* rb = ra + NA * n;
        double Dbd[NB * ND], Dbc[NB * NC], Dad[NA * ND], Dac[NA * NC];
#pragma unroll
        for (int i = 0; i < NB; ++i) {
#pragma unroll
            for (int j = 0; j < ND; ++j) Dbd[i * ND + j] = D[(size_t)(b0 + i) * n + d0 + j];
#pragma unroll
            for (int j = 0; j < NC; ++j) Dbc[i * NC + j] = D[(size_t)(b0 + i) * n + c0 + j];
        }
#pragma unroll
        for (int i = 0; i < NA; ++i) {
#pragma unroll
            for (int j = 0; j < ND; ++j) Dad[i * ND + j] = D[(size_t)(a0 + i) * n + d0 + j];
#pragma unroll
            for (int j = 0; j < NC; ++j) Dac[i * NC + j] = D[(size_t)(a0 + i) * n + c0 + j];
        }
#pragma unroll
        for (int ia = 0; ia < NA; ++ia) {
#pragma unroll
            for (int ic = 0; ic < NC; ++ic) {
                double acc = 0;
#pragma unroll
                for (int ib = 0; ib < NB; ++ib)
#pragma unroll
                    for (int id = 0; id < ND; ++id) acc += v[((ia * NB + ib) * NC + ic) * ND + id] * Dbd[ib * ND + id];
                atomicAdd(&ra[ia * n + c0 + ic], w * acc);
            }
#pragma unroll
            for (int id = 0; id < ND; ++id) {
                double acc = 0;
#pragma unroll
                for (int ib = 0; ib < NB; ++ib)
#pragma unroll
                    for (int ic = 0; ic < NC; ++ic) acc += v[((ia * NB + ib) * NC + ic) * ND + id] * Dbc[ib * NC + ic];
                atomicAdd(&ra[ia * n + d0 + id], w * acc);
            }
        }
#pragma unroll
        for (int ib = 0; ib < NB; ++ib) {
#pragma unroll
            for (int ic = 0; ic < NC; ++ic) {
                double acc = 0;
#pragma unroll
                for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                    for (int id = 0; id < ND; ++id) acc += v[((ia * NB + ib) * NC + ic) * ND + id] * Dad[ia * ND + id];
                atomicAdd(&rb[ib * n + c0 + ic], w * acc);
            }
#pragma unroll
            for (int id = 0; id < ND; ++id) {
                double acc = 0;
#pragma unroll
                for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                    for (int ic = 0; ic < NC; ++ic) acc += v[((ia * NB + ib) * NC + ic) * ND + id] * Dac[ia * NC + ic];
                atomicAdd(&rb[ib * n + d0 + id], w * acc);
            }
        }
    }
}

template <class C>
__device__ __forceinline__ void sink_packed(const double* __restrict__ v, int A, int B, int Cs, int Ds, int a0, int b0,
                                            int c0, int d0, bool same_pair, const SinkArgs& s) {
    constexpr int NA = C::NA, NB = C::NB, NC = C::NC, ND = C::ND;
#pragma unroll
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
        for (int ib = 0; ib < NB; ++ib) {
            const int a = a0 + ia, b = b0 + ib;
            if (A == B && a < b) continue;
            const uint64_t ab = two_index(a, b);
#pragma unroll
            for (int ic = 0; ic < NC; ++ic)
#pragma unroll
                for (int id = 0; id < ND; ++id) {
                    const int c = c0 + ic, d = d0 + id;
                    if (Cs == Ds && c < d) continue;
                    const uint64_t cd = two_index(c, d);
                    if (same_pair && ab < cd) continue;
                    const double x = v[((ia * NB + ib) * NC + ic) * ND + id];
                    if (s.pair_stride) {
                        const uint64_t pab = (uint64_t)pair_pad_index(a, b), pcd = (uint64_t)pair_pad_index(c, d);
                        const uint64_t i1 = pab * s.pair_stride + pcd, i2 = pcd * s.pair_stride + pab;
                        if (i1 < s.packed_len && i2 < s.packed_len) {
                            s.packed[i1] = x;
                            s.packed[i2] = x;
                        }
                    } else {
                        const uint64_t idx = two_index(ab, cd);
                        if (idx < s.packed_len) s.packed[idx] = x;
                    }
                }
        }
}

// One contracted shell quartet: primitive loops (Boys, VRR, transfer, contraction) then HRR and the
// component normalisation.  A..D are the four centres; out receives NA*NB*NC*ND values, index
// ((ia*NB + ib)*NC + ic)*ND + id.
template <class C, typename R = double, bool BOYS_LDS = false>
__host__ __device__ __forceinline__ void eval_quartet(const double* __restrict__ A, const double* __restrict__ B,
                                                     const double* __restrict__ Cc, const double* __restrict__ Dd,
                                                     const PrimPair* __restrict__ bra, int npb,
                                                     const PrimPair* __restrict__ ket, int npk,
                                                     const R* __restrict__ boys, R* __restrict__ out) {
    R acc[C::NACC];
#pragma unroll
    for (int i = 0; i < C::NACC; ++i) acc[i] = R(0);
    auto prim = [&](const PrimPair& P, const PrimPair& Q) __attribute__((always_inline)) {
        const R PAx = R(P.Px - A[0]), PAy = R(P.Py - A[1]), PAz = R(P.Pz - A[2]);
        const R zeta = R(P.zeta), inv2z = R(0.5 * P.inv_zeta);
        const R eta = R(Q.zeta), inv_eta = R(Q.inv_zeta), inv2e = R(0.5) * inv_eta;
        const R rs = rsqrt_r(zeta + eta);               // (zeta + eta)^{-1/2}
        const R inv_s = rs * rs;
        const R rho = zeta * eta * inv_s;
        const R PQx = R(P.Px - Q.Px), PQy = R(P.Py - Q.Py), PQz = R(P.Pz - Q.Pz);
        const R T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
        const R pref = R(kTwoPi25) * R(P.Kz) * R(Q.Kz) * rs;
        R F[C::L + 1];
        if constexpr (sizeof(R) == sizeof(double) && BOYS_LDS)
            boys_scaled<C::L, kBoysLdsStride, 0, 8>(T, pref, boys, F);
        else if constexpr (sizeof(R) == sizeof(double))
            boys_scaled<C::L>(T, pref, boys, F);
        else if constexpr (BOYS_LDS)
            boys_scaled_f32<C::L, kBoysLdsStride32, 0>(T, pref, boys, F);
        else
            boys_scaled_f32<C::L>(T, pref, boys, F);
        const R rz = eta * inv_s;                       // rho / zeta
        const R zoe = zeta * inv_eta;                   // zeta / eta
        const R dlx = zoe * PAx + R(Q.Px - Cc[0]);      // delta = (zeta/eta) PA + QC
        const R dly = zoe * PAy + R(Q.Py - Cc[1]);
        const R dlz = zoe * PAz + R(Q.Pz - Cc[2]);
        C::prim(F, PAx, PAy, PAz, -rz * PQx, -rz * PQy, -rz * PQz, inv2z, rz, dlx, dly, dlz, zoe, inv2e, acc);
    };
    // The loop is bound by load latency (primitive pair, then the Boys row it selects).  The lanes of
    // a wavefront hold different kets but (tasks being ordered by bra pair) mostly the same bra, so
    // ket primitives go in the outer loop: each lane loads each of its ket primitive pairs once, and
    // the inner loop's bra loads are the same address across the wavefront.  Small classes take the
    // inner primitives two at a time so both chains are in flight together; a one-primitive bra
    // (d shells, outer valence shells) pairs the ket primitives instead.
    constexpr bool kPairs = C::NACC <= kKetPairMaxAcc;
    if (kPairs && npb == 1) {
        const PrimPair P = bra[0];
        int j = 0;
        for (; j + 1 < npk; j += 2) {
            const PrimPair Q0 = ket[j], Q1 = ket[j + 1];
            prim(P, Q0);
            prim(P, Q1);
        }
        if (j < npk) prim(P, ket[j]);
    } else {
        for (int j = 0; j < npk; ++j) {
            const PrimPair Q = ket[j];
            int i = 0;
            if constexpr (kPairs) {
                for (; i + 1 < npb; i += 2) {
                    const PrimPair P0 = bra[i], P1 = bra[i + 1];
                    prim(P0, Q);
                    prim(P1, Q);
                }
            }
            for (; i < npb; ++i) prim(bra[i], Q);
        }
    }
    C::hrr(acc, R(A[0] - B[0]), R(A[1] - B[1]), R(A[2] - B[2]), R(Cc[0] - Dd[0]), R(Cc[1] - Dd[1]), R(Cc[2] - Dd[2]),
           out);
}

// A class whose reference convention truncates the transfer step (C::TRUNCATED) has values that
// depend on which pair is the bra.  The reference chooses per component
// (HartreeFock/IntegralsRepository.cpp:344-370, 436-475): after ordering each pair by basis-function
// ID, the bra is the pair with the larger (first ID, second ID).  Pairs are stored with the higher
// shell index first, so shell indices decide unless both pairs start with the same shell; then the
// block is evaluated both ways and each component takes its own choice.  Bra and ket classes are
// equal for every truncated class (L <= 2: only (dd|dd)).
template <class C>
__host__ __device__ __forceinline__ void eval_quartet_reference_order(
    int A, int Cs, const double* __restrict__ xa, const double* __restrict__ xb, const double* __restrict__ xc,
    const double* __restrict__ xd, const PrimPair* __restrict__ bra, int npb, const PrimPair* __restrict__ ket, int npk,
    int a0, int b0, int c0, int d0, const double* __restrict__ boys, double* __restrict__ out) {
    static_assert(C::NA == C::NC && C::NB == C::ND, "truncated classes have equal bra and ket classes");
    constexpr int NA = C::NA, NB = C::NB, NC = C::NC, ND = C::ND;
    const bool needP = A >= Cs, needQ = A <= Cs;
    if (needP) eval_quartet<C>(xa, xb, xc, xd, bra, npb, ket, npk, boys, out);
    if (!needQ) return;
    double t[C::NOUT];
    eval_quartet<C>(xc, xd, xa, xb, ket, npk, bra, npb, boys, t);     // t[c][d][a][b]
    for (int ia = 0; ia < NA; ++ia)
        for (int ib = 0; ib < NB; ++ib)
            for (int ic = 0; ic < NC; ++ic)
                for (int id = 0; id < ND; ++id) {
                    const int a = a0 + ia, b = b0 + ib, c = c0 + ic, d = d0 + id;
                    const int i = a > b ? a : b, j = a > b ? b : a, k = c > d ? c : d, l = c > d ? d : c;
                    const bool braP = needP && !(i < k || (i == k && j < l));
                    if (!braP) out[((ia * NB + ib) * NC + ic) * ND + id] = t[((ic * ND + id) * NA + ia) * NB + ib];
                }
}

// sl: primitive slice + 1 of a sliced J/K task (0: the whole quartet); target: the class's slice size.
template <class C, typename R = double, bool BOYS_LDS = false>
__device__ __forceinline__ void quartet_values(const ShellsDev& sh, const PrimPair* __restrict__ pp,
                                               const R* __restrict__ boys, const PairRec& bp, const PairRec& kp,
                                               R* out, int sl = 0, int target = 0) {
    const double *xa = sh.xyz + 3 * bp.sa, *xb = sh.xyz + 3 * bp.sb, *xc = sh.xyz + 3 * kp.sa, *xd = sh.xyz + 3 * kp.sb;
    if constexpr (C::TRUNCATED) {
        static_assert(sizeof(R) == sizeof(double), "the reference-convention classes are FP64 only");
        eval_quartet_reference_order<C>(bp.sa, kp.sa, xa, xb, xc, xd, pp + bp.prim0, bp.nprim, pp + kp.prim0, kp.nprim,
                                        bp.a0, bp.b0, kp.a0, kp.b0, boys, out);
    } else {
        int b_off = 0, nb = bp.nprim;
        if (sl > 0) {
            const int chunk = slice_chunk(bp.nprim, kp.nprim, target);
            b_off = (sl - 1) * chunk;
            nb = min(chunk, bp.nprim - b_off);
        }
        eval_quartet<C, R, BOYS_LDS>(xa, xb, xc, xd, pp + bp.prim0 + b_off, nb, pp + kp.prim0, kp.nprim, boys, out);
    }
}

// J/K digestion for a block of consecutive tasks.  Tasks are ordered by bra pair, so a block covers
// a short run of bra pairs [pb, pe].  The first sink.lds_slots of them get an LDS slot holding
// their exchange rows; every K' contribution of their quartets is an LDS atomic, and at the end
// the block adds the slots to K' with column-coalesced global atomics (one per touched element
// instead of one per quartet and component).  J'_ab is summed across the wavefront when the
// wavefront shares one slot.  Quartets of further bra pairs use the plain global-atomic sink.
template <class C, typename R>
__device__ __forceinline__ void eri_jk_block(const TaskList& tl, const ShellsDev& sh, const PrimPair* __restrict__ pp,
                                             const R* __restrict__ boys, const SinkArgs& sink) {
    extern __shared__ double rows[];
    constexpr int NA = C::NA, NB = C::NB, NAB = C::NA * C::NB;
    const int n = sink.n;
    const size_t nn = (size_t)n * n;
    const int rows_per_slot = sink.ndens * (NA + NB);
    const int slot_len = rows_per_slot * n;
    __shared__ int s_pe;
    // this workgroup's task blocks: tasks_per_block consecutive 256-task blocks (the last group may be shorter)
    const int tpb = sink.tasks_per_block > 1 ? sink.tasks_per_block : 1;
    const int64_t blk0 = (int64_t)blockIdx.x * tpb;
    const int nblk = (int)(tl.nblocks - blk0 < tpb ? tl.nblocks - blk0 : tpb);
    const int pb = tl.tasks[(size_t)blk0 * blockDim.x].x;     // a block's first task is never padding
    if (threadIdx.x == 0) s_pe = pb;
    __syncthreads();
    for (int r = 0; r < nblk; ++r) {
        const int x = tl.tasks[(size_t)(blk0 + r) * blockDim.x + threadIdx.x].x;
        if (x > pb) atomicMax(&s_pe, x);
    }
    __syncthreads();
    const int pe = s_pe;
    const int nslot = pe - pb + 1 < sink.lds_slots ? pe - pb + 1 : sink.lds_slots;
    for (int i = threadIdx.x; i < nslot * slot_len; i += blockDim.x) rows[i] = 0.0;
    constexpr bool kLdsBoys = boys_in_lds<C>();
    R* btab = (R*)(rows + (size_t)sink.lds_slots * slot_len);
    if constexpr (kLdsBoys && sizeof(R) == sizeof(double)) {
        for (int i = threadIdx.x; i < BOYS_NROW * kBoysLdsStride; i += blockDim.x) {
            const int r = i / kBoysLdsStride, c = i - r * kBoysLdsStride;
            btab[i] = boys[r * BOYS_STRIDE + (c < 8 ? C::L + c : BOYS_NCOL)];
        }
    } else if constexpr (kLdsBoys) {
        for (int i = threadIdx.x; i < BOYS_NROW * kBoysLdsStride32; i += blockDim.x) {
            const int r = i / kBoysLdsStride32, c = i - r * kBoysLdsStride32;
            btab[i] = c < 6 ? boys[r * BOYS_STRIDE + C::L + c] : R(0);
        }
    }
    __syncthreads();

    for (int r = 0; r < nblk; ++r) {
        const int2 me = tl.tasks[(size_t)(blk0 + r) * blockDim.x + threadIdx.x];
        const bool active = me.x >= 0;
        int key = -1, a0 = 0, b0 = 0;
        double jab[NAB];
#pragma unroll
        for (int i = 0; i < NAB; ++i) jab[i] = 0.0;
        if (active) {
            const int p = me.x, q = me.y & kTaskKetMask, sl = me.y >> kTaskKetBits;
            const PairRec bp = tl.bra.rec[p], kp = tl.ket.rec[q];
            const int A = bp.sa, B = bp.sb, Cs = kp.sa, Ds = kp.sb;
            a0 = bp.a0;
            b0 = bp.b0;
            const int c0 = kp.a0, d0 = kp.b0;
            R out[C::NOUT];
            if constexpr (kLdsBoys)
                quartet_values<C, R, true>(sh, pp, (const R*)btab, bp, kp, out, sl, tl.slice_prims);
            else
                quartet_values<C, R>(sh, pp, boys, bp, kp, out, sl, tl.slice_prims);
            double w = 1.0;
            if (A == B) w *= 0.5;
            if (Cs == Ds) w *= 0.5;
            if (tl.same_class && p == q) w *= 0.5;
            // (a later block of the group may start a new segment, whose bra pairs can precede pb)
            if (p >= pb && p - pb < nslot) {
                key = p - pb;
                sink_jk_rows<C, R>(out, rows + (size_t)key * slot_len, a0, b0, c0, d0, w, sink, jab);
            } else {
                sink_jk<C, R>(out, a0, b0, c0, d0, w, sink);
            }
        }
        // J'_ab: lanes without a slot hold zeros; lane 0 is active whenever any lane is
        const int key0 = __shfl(key, 0);
        if (__all(key == key0 || key < 0)) {
            // the NAB sums by recursive halving in registers (padded to a power of two); sum v ends in lane
            // v << SH, and those lanes add their sums to J' in one atomic instruction
            static_assert(NAB <= 64, "one value per lane at most");
            constexpr int CNT = NAB <= 1 ? 1 : NAB <= 2 ? 2 : NAB <= 4 ? 4 : NAB <= 8 ? 8 : NAB <= 16 ? 16 : NAB <= 32 ? 32 : 64;
            constexpr int SH = 6 - log2c<CNT>();
            double V[CNT];
#pragma unroll
            for (int i = 0; i < CNT; ++i) V[i] = i < NAB ? jab[i] : 0.0;
            const int lane = threadIdx.x & 63;
            Halving<CNT, 32>::run(V, lane);
            const int v = lane >> SH;
            const int ra0 = __builtin_amdgcn_readlane(a0, 0), rb0 = __builtin_amdgcn_readlane(b0, 0);
            if (key0 >= 0 && (lane & ((1 << SH) - 1)) == 0 && v < NAB)
                unsafeAtomicAdd(&sink.J[(size_t)(ra0 + v / NB) * n + rb0 + v % NB], V[0]);
        } else if (key >= 0) {
#pragma unroll
            for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                for (int ib = 0; ib < NB; ++ib) unsafeAtomicAdd(&sink.J[(size_t)(a0 + ia) * n + b0 + ib], jab[ia * NB + ib]);
        }
    }
    __syncthreads();
    // flush: row by row, consecutive threads on consecutive columns
    for (int sr = 0; sr < nslot * rows_per_slot; ++sr) {
        const int s = sr / rows_per_slot, kr = sr - s * rows_per_slot;
        const int k = kr / (NA + NB), r = kr - k * (NA + NB);
        const int p = pb + s;
        const int grow = r < NA ? tl.bra.rec[p].a0 + r : tl.bra.rec[p].b0 + (r - NA);
        const double* src = rows + (size_t)sr * n;
        double* dst = sink.K + k * nn + (size_t)grow * n;
        for (int col = threadIdx.x; col < n; col += blockDim.x) {
            const double x = src[col];
            if (x != 0.0) unsafeAtomicAdd(&dst[col], x);
        }
    }
}

template <int LA, int LB, int LC, int LD, int EXACT, int MODE>
__global__ void __launch_bounds__(256) eri_class_kernel(TaskList tl, ShellsDev sh, const PrimPair* __restrict__ pp,
                                                        const double* __restrict__ boys, SinkArgs sink) {
    using C = EriClass<LA, LB, LC, LD, EXACT>;
    if constexpr (MODE == MODE_JK) {
        eri_jk_block<C, double>(tl, sh, pp, boys, sink);
    } else if constexpr (MODE == MODE_JK32) {
        eri_jk_block<C, float>(tl, sh, pp, sink.boys32, sink);
    } else {
        // The Boys rows of this class (F_L..F_{L+7} and e^{-T_k}) staged in LDS, as the J/K kernels do: the table
        // row a primitive quartet selects is then an LDS load, not an L2 round trip, in the loop's dependent
        // chain (T, row, polynomial).  Every thread stages before any leaves (padding tasks, the grid's tail).
        constexpr bool kLdsBoys = boys_in_lds<C>();
        extern __shared__ double btab[];
        if constexpr (kLdsBoys) {
            for (int i = threadIdx.x; i < BOYS_NROW * kBoysLdsStride; i += blockDim.x) {
                const int r = i / kBoysLdsStride, c = i - r * kBoysLdsStride;
                btab[i] = boys[r * BOYS_STRIDE + (c < 8 ? C::L + c : BOYS_NCOL)];
            }
            __syncthreads();
        }
        int p, q;
        if (MODE == MODE_DIAG) {
            const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
            if (t >= tl.ntasks) return;
            p = q = (int)t;
        } else {
            const int2 me = tl.tasks[(size_t)blockIdx.x * blockDim.x + threadIdx.x];
            if (me.x < 0) return;
            p = me.x;
            q = me.y & kTaskKetMask;
        }
        const PairRec bp = tl.bra.rec[p], kp = tl.ket.rec[q];
        const int A = bp.sa, B = bp.sb, Cs = kp.sa, Ds = kp.sb;
        const int a0 = bp.a0, b0 = bp.b0, c0 = kp.a0, d0 = kp.b0;
        double out[C::NOUT];
        if constexpr (kLdsBoys)
            quartet_values<C, double, true>(sh, pp, (const double*)btab, bp, kp, out);
        else
            quartet_values<C>(sh, pp, boys, bp, kp, out);
        if constexpr (MODE == MODE_PACKED) {
            sink_packed<C>(out, A, B, Cs, Ds, a0, b0, c0, d0, tl.same_class && p == q, sink);
        } else {
            double mx = 0.0;
#pragma unroll
            for (int ia = 0; ia < C::NA; ++ia)
#pragma unroll
                for (int ib = 0; ib < C::NB; ++ib) {
                    const double x = fabs(out[((ia * C::NB + ib) * C::NC + ia) * C::ND + ib]);
                    mx = x > mx ? x : mx;
                }
            sink.Qout[p] = sqrt(mx);
        }
    }
}

template <int LA, int LB, int LC, int LD, int EXACT, int MODE>
inline hipError_t launch_eri(const TaskList& tl, const ShellsDev& sh, const PrimPair* pp, const double* boys,
                             const SinkArgs& sink, hipStream_t st) {
    if (tl.ntasks <= 0) return hipSuccess;
    int64_t blocks = MODE == MODE_DIAG ? (tl.ntasks + 255) / 256 : tl.nblocks;
    SinkArgs s = sink;
    size_t lds = 0;
    if (MODE == MODE_JK || MODE == MODE_JK32) {
        using C = EriClass<LA, LB, LC, LD, 0>;
        // A block flushes its exchange rows (ndens (NA + NB) rows of n columns) with one coalesced atomic per
        // element; where that is more than 8 atomics per quartet (long rows: large n, two densities) a workgroup
        // digests four 256-task blocks before its flush, when that still leaves >= 1024 workgroups.  (For short
        // rows the fewer, longer workgroups overlap less with the build's other classes and cost more than the
        // flushes they save.)
        const int64_t flush = (int64_t)sink.ndens * (C::NA + C::NB) * sink.n;
        s.tasks_per_block = (flush > 8 * 256 && blocks >= 4096) ? 4 : 1;
        blocks = (blocks + s.tasks_per_block - 1) / s.tasks_per_block;
        // up to 4 bra pairs per block, within 40 KB of LDS (a block of 256 tasks rarely spans more)
        const size_t slot_bytes = (size_t)sink.ndens * (C::NA + C::NB) * sink.n * sizeof(double);
        const size_t cap = 40 * 1024;
        // (none in deterministic mode: the slots' LDS atomics add in no fixed order)
        s.lds_slots = sink.fx ? 0 : (int)std::min<size_t>(4, cap / slot_bytes);
        lds = s.lds_slots * slot_bytes;
        if (boys_in_lds<C>()) lds += kBoysLdsBytes;    // (the FP32 slice needs less)
    } else if (boys_in_lds<EriClass<LA, LB, LC, LD, EXACT>>()) {
        lds = kBoysLdsBytes;                           // MODE_PACKED, MODE_DIAG: the Boys rows in LDS
    }
    hipLaunchKernelGGL((eri_class_kernel<LA, LB, LC, LD, EXACT, MODE>), dim3((unsigned)blocks), dim3(256), lds, st, tl,
                       sh, pp, boys, s);
    return hipGetLastError();
}

}  // namespace hfgpu
