// Cooperative shell-quartet kernel for the high angular-momentum classes: one wavefront per
// contracted quartet.  The recursion intermediates of the quartet live in LDS and the 64 lanes
// execute each dependency level of the generated op list in parallel:
//     val[dst] = sum_{r<4} coef[c_r] * val[s_r]
// (vertical recursion and electron transfer per primitive quartet, accumulation into the
// contracted block, then the horizontal recursions).  A thread-per-quartet kernel would need
// ~1-30 KB of scratch per lane for these classes; here the state is ~6-49 KB of LDS per wave.
// Sinks are the same three as eri_kernel.hpp (J/K digestion, packed store, Schwarz diagonal),
// with their reductions spread over the lanes.
#pragma once

#include "eri_kernel.hpp"

namespace hfgpu {

struct CoopOp {
    unsigned short dst;
    unsigned short s[4];
    unsigned char c[4];
};

template <int LA, int LB, int LC, int LD, int EXACT>
struct CoopClass;

constexpr int kCoopCoef = 48;   // coefficient ids 0..41 used (see gen_eri.py COEF)

template <class K>
__device__ __forceinline__ void coop_run(const CoopOp* __restrict__ ops, int begin, int end, double* __restrict__ val,
                                         const double* __restrict__ cf, int lane) {
    for (int k = begin + lane; k < end; k += 64) {
        const CoopOp o = ops[k];
        val[o.dst] = cf[o.c[0]] * val[o.s[0]] + cf[o.c[1]] * val[o.s[1]] + cf[o.c[2]] * val[o.s[2]] +
                     cf[o.c[3]] * val[o.s[3]];
    }
}

// Evaluates one contracted quartet (bra pair primitives `bra`, ket pair primitives `ket`); the
// unnormalised block ends in val[K::out_slot()[idx]].  All lanes of the wave call it.
template <class K>
__device__ void coop_evaluate(double* __restrict__ val, double* __restrict__ cf, int lane, const double* xa,
                              const double* xb, const double* xc, const double* xd,
                              const PrimPair* __restrict__ bra, int npb, const PrimPair* __restrict__ ket, int npk,
                              const double* __restrict__ boys) {
    for (int i = lane; i < K::ACC0 + K::NACC; i += 64) val[i] = 0.0;
    if (lane == 0) {
        cf[0] = 0.0;
        cf[1] = 1.0;
    }
    for (int i = 0; i < npb; ++i)
        for (int j = 0; j < npk; ++j) {
            if (lane == 0) {
                const PrimPair P = bra[i];
                const PrimPair Q = ket[j];
                const double PAx = P.Px - xa[0], PAy = P.Py - xa[1], PAz = P.Pz - xa[2];
                const double inv_s = 1.0 / (P.zeta + Q.zeta);
                const double rho = P.zeta * Q.zeta * inv_s;
                const double PQx = P.Px - Q.Px, PQy = P.Py - Q.Py, PQz = P.Pz - Q.Pz;
                const double T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
                const double pref = kTwoPi25 * P.Kz * Q.Kz * sqrt(inv_s);
                double F[K::L + 1];
                boys_scaled<K::L>(T, pref, boys, F);
#pragma unroll
                for (int m = 0; m <= K::L; ++m) val[K::F0 + m] = F[m];
                const double rz = Q.zeta * inv_s;
                const double zoe = P.zeta * Q.inv_zeta;
                const double inv2z = 0.5 * P.inv_zeta, inv2e = 0.5 * Q.inv_zeta;
                cf[2] = PAx; cf[3] = PAy; cf[4] = PAz;
                cf[5] = -rz * PQx; cf[6] = -rz * PQy; cf[7] = -rz * PQz;
                cf[8] = zoe * PAx + (Q.Px - xc[0]);
                cf[9] = zoe * PAy + (Q.Py - xc[1]);
                cf[10] = zoe * PAz + (Q.Pz - xc[2]);
                cf[11] = -zoe;
#pragma unroll
                for (int k = 1; k <= 8; ++k) {
                    cf[11 + k] = k * inv2z;
                    cf[19 + k] = -rz * k * inv2z;
                    cf[27 + k] = k * inv2e;
                }
            }
            __syncthreads();
            const CoopOp* ops = K::prim_ops();
#pragma unroll 1
            for (int l = 0; l < K::NPL; ++l) {
                coop_run<K>(ops, K::PL[l], K::PL[l + 1], val, cf, lane);
                __syncthreads();
            }
        }
    if (lane == 0) {
        cf[36] = xa[0] - xb[0]; cf[37] = xa[1] - xb[1]; cf[38] = xa[2] - xb[2];
        cf[39] = xc[0] - xd[0]; cf[40] = xc[1] - xd[1]; cf[41] = xc[2] - xd[2];
    }
    __syncthreads();
    const CoopOp* ops = K::hrr_ops();
#pragma unroll 1
    for (int l = 0; l < K::NHL; ++l) {
        coop_run<K>(ops, K::HL[l], K::HL[l + 1], val, cf, lane);
        __syncthreads();
    }
}

template <class K, int MODE>
__global__ void __launch_bounds__(64) eri_coop_kernel(TaskList tl, ShellsDev sh, const PrimPair* __restrict__ pp,
                                                      const double* __restrict__ boys, SinkArgs sink) {
    extern __shared__ double val[];
    __shared__ double cf[kCoopCoef];
    const int lane = threadIdx.x;
    int p, q;
    if (MODE == MODE_DIAG) {
        if ((int64_t)blockIdx.x >= tl.ntasks) return;
        p = q = (int)blockIdx.x;
    } else {
        const int2 me = tl.tasks[blockIdx.x];
        if (me.x < 0) return;                      // segment padding
        p = me.x;
        q = me.y & kTaskKetMask;                   // (these classes are never sliced)
    }
    const PairRec bp = tl.bra.rec[p], kp = tl.ket.rec[q];
    const int A = bp.sa, B = bp.sb, Cs = kp.sa, Ds = kp.sb;
    const double* xa = sh.xyz + 3 * A;
    const double* xb = sh.xyz + 3 * B;
    const double* xc = sh.xyz + 3 * Cs;
    const double* xd = sh.xyz + 3 * Ds;
    const PrimPair* bra = pp + bp.prim0;
    const PrimPair* ket = pp + kp.prim0;
    const int npb = bp.nprim, npk = kp.nprim;
    const int a0 = bp.a0, b0 = bp.b0, c0 = kp.a0, d0 = kp.b0;
    constexpr int NA = K::NA, NB = K::NB, NC = K::NC, ND = K::ND;
    const unsigned short* os = K::out_slot();
    const double* on = K::out_norm();
    double* res = val + K::NSLOT;     // truncated classes: the selected block, normalised
    if constexpr (K::TRUNCATED) {
        // reference convention: per component, the bra is the pair with the larger (first ID,
        // second ID) (HartreeFock/IntegralsRepository.cpp:344-370); pairs are stored with the
        // higher shell first, so the shell indices decide unless both pairs start with one shell.
        const bool needP = A >= Cs, needQ = A <= Cs;
        if (needP) {
            coop_evaluate<K>(val, cf, lane, xa, xb, xc, xd, bra, npb, ket, npk, boys);
            for (int idx = lane; idx < K::NOUT; idx += 64) res[idx] = val[os[idx]] * on[idx];
            __syncthreads();
        }
        if (needQ) {
            coop_evaluate<K>(val, cf, lane, xc, xd, xa, xb, ket, npk, bra, npb, boys);   // (Q|P) block
            for (int idx = lane; idx < K::NOUT; idx += 64) {
                const int id = idx % ND, ic = (idx / ND) % NC, ib = (idx / (ND * NC)) % NB, ia = idx / (ND * NC * NB);
                const int a = a0 + ia, b = b0 + ib, c = c0 + ic, d = d0 + id;
                const int i = a > b ? a : b, j = a > b ? b : a, k = c > d ? c : d, l = c > d ? d : c;
                const bool braP = needP && !(i < k || (i == k && j < l));
                const int tidx = ((ic * ND + id) * NA + ia) * NB + ib;
                if (!braP) res[idx] = val[os[tidx]] * on[tidx];
            }
            __syncthreads();
        }
    } else {
        coop_evaluate<K>(val, cf, lane, xa, xb, xc, xd, bra, npb, ket, npk, boys);
    }
    // final normalised block value v(ia,ib,ic,id) at index ((ia*NB+ib)*NC+ic)*ND+id
    auto v = [&](int idx) { return K::TRUNCATED ? res[idx] : val[os[idx]] * on[idx]; };
    if (MODE == MODE_DIAG) {
        if (lane == 0) {
            double mx = 0.0;
            for (int ia = 0; ia < NA; ++ia)
                for (int ib = 0; ib < NB; ++ib) {
                    const double x = fabs(v(((ia * NB + ib) * NC + ia) * ND + ib));
                    mx = x > mx ? x : mx;
                }
            sink.Qout[p] = sqrt(mx);
        }
    } else if (MODE == MODE_PACKED) {
        const bool same_pair = tl.same_class && p == q;
        for (int idx = lane; idx < K::NOUT; idx += 64) {
            const int id = idx % ND, ic = (idx / ND) % NC, ib = (idx / (ND * NC)) % NB, ia = idx / (ND * NC * NB);
            const int a = a0 + ia, b = b0 + ib, c = c0 + ic, d = d0 + id;
            if (A == B && a < b) continue;
            if (Cs == Ds && c < d) continue;
            const uint64_t ab = two_index(a, b), cd = two_index(c, d);
            if (same_pair && ab < cd) continue;
            if (sink.pair_stride) {
                const uint64_t pab = (uint64_t)pair_pad_index(a, b), pcd = (uint64_t)pair_pad_index(c, d);
                const uint64_t i1 = pab * sink.pair_stride + pcd, i2 = pcd * sink.pair_stride + pab;
                if (i1 < sink.packed_len && i2 < sink.packed_len) {
                    const double x = v(idx);
                    sink.packed[i1] = x;
                    sink.packed[i2] = x;
                }
            } else {
                const uint64_t pidx = two_index(ab, cd);
                if (pidx < sink.packed_len) sink.packed[pidx] = v(idx);
            }
        }
    } else {
        double w = 1.0;
        if (A == B) w *= 0.5;
        if (Cs == Ds) w *= 0.5;
        if (tl.same_class && p == q) w *= 0.5;
        const int n = sink.n;
        const size_t nn = (size_t)n * n;
        const double w2 = 2.0 * w;
        // Coulomb with the summed densities
        for (int r = lane; r < NA * NB + NC * ND; r += 64) {
            double acc = 0;
            if (r < NA * NB) {
                const int ia = r / NB, ib = r % NB;
                for (int ic = 0; ic < NC; ++ic)
                    for (int id = 0; id < ND; ++id) {
                        double dsum = 0;
                        for (int k = 0; k < sink.ndens; ++k) dsum += sink.D[k * nn + (size_t)(c0 + ic) * n + d0 + id];
                        acc += v(((ia * NB + ib) * NC + ic) * ND + id) * dsum;
                    }
                sink_add_J(sink, (size_t)(a0 + ia) * n + b0 + ib, w2 * acc);
            } else {
                const int rr = r - NA * NB, ic = rr / ND, id = rr % ND;
                for (int ia = 0; ia < NA; ++ia)
                    for (int ib = 0; ib < NB; ++ib) {
                        double dsum = 0;
                        for (int k = 0; k < sink.ndens; ++k) dsum += sink.D[k * nn + (size_t)(a0 + ia) * n + b0 + ib];
                        acc += v(((ia * NB + ib) * NC + ic) * ND + id) * dsum;
                    }
                sink_add_J(sink, (size_t)(c0 + ic) * n + d0 + id, w2 * acc);
            }
        }
        // Exchange per density: K'_ac, K'_ad, K'_bc, K'_bd
        for (int k = 0; k < sink.ndens; ++k) {
            const double* __restrict__ Dm = sink.D + k * nn;
            constexpr int R1 = NA * NC, R2 = R1 + NA * ND, R3 = R2 + NB * NC, R4 = R3 + NB * ND;
            for (int r = lane; r < R4; r += 64) {
                double acc = 0;
                if (r < R1) {
                    const int ia = r / NC, ic = r % NC;
                    for (int ib = 0; ib < NB; ++ib)
                        for (int id = 0; id < ND; ++id)
                            acc += v(((ia * NB + ib) * NC + ic) * ND + id) * Dm[(size_t)(b0 + ib) * n + d0 + id];
                    sink_add_K(sink, k, (size_t)(a0 + ia) * n + c0 + ic, w * acc);
                } else if (r < R2) {
                    const int rr = r - R1, ia = rr / ND, id = rr % ND;
                    for (int ib = 0; ib < NB; ++ib)
                        for (int ic = 0; ic < NC; ++ic)
                            acc += v(((ia * NB + ib) * NC + ic) * ND + id) * Dm[(size_t)(b0 + ib) * n + c0 + ic];
                    sink_add_K(sink, k, (size_t)(a0 + ia) * n + d0 + id, w * acc);
                } else if (r < R3) {
                    const int rr = r - R2, ib = rr / NC, ic = rr % NC;
                    for (int ia = 0; ia < NA; ++ia)
                        for (int id = 0; id < ND; ++id)
                            acc += v(((ia * NB + ib) * NC + ic) * ND + id) * Dm[(size_t)(a0 + ia) * n + d0 + id];
                    sink_add_K(sink, k, (size_t)(b0 + ib) * n + c0 + ic, w * acc);
                } else {
                    const int rr = r - R3, ib = rr / ND, id = rr % ND;
                    for (int ia = 0; ia < NA; ++ia)
                        for (int ic = 0; ic < NC; ++ic)
                            acc += v(((ia * NB + ib) * NC + ic) * ND + id) * Dm[(size_t)(a0 + ia) * n + c0 + ic];
                    sink_add_K(sink, k, (size_t)(b0 + ib) * n + d0 + id, w * acc);
                }
            }
        }
    }
}

template <class K, int MODE>
inline hipError_t launch_coop(const TaskList& tl, const ShellsDev& sh, const PrimPair* pp, const double* boys,
                              const SinkArgs& sink, hipStream_t st) {
    if (tl.ntasks <= 0) return hipSuccess;
    const size_t lds = (K::NSLOT + (K::TRUNCATED ? K::NOUT : 0)) * sizeof(double);
    // one block per entry of the expanded task list (padding entries return at once)
    const int64_t blocks = MODE == MODE_DIAG ? tl.ntasks : tl.nblocks * 256;
    hipLaunchKernelGGL((eri_coop_kernel<K, MODE>), dim3((unsigned)blocks), dim3(64), lds, st, tl, sh, pp, boys,
                       sink);
    return hipGetLastError();
}

}  // namespace hfgpu
