// Wavefront reductions of gfx950 in registers: lane exchanges by v_permlane32_swap / v_permlane16_swap
// and DPP moves instead of ds_bpermute (an LDS round trip per exchange).
#pragma once
#include <hip/hip_runtime.h>

namespace hfgpu {

// Lane exchanges of gfx950 that stay in registers (no LDS round trip).  swap32 / swap16 are
// v_permlane32_swap / v_permlane16_swap on both halves of a double: lanes 32..63 of a trade places with
// lanes 0..31 of b (swap32), or the odd 16-lane rows of a with the even rows of b (swap16).
__device__ __forceinline__ double join_halves(unsigned lo, unsigned hi) {
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <int S>
__device__ __forceinline__ void lane_swap(double& a, double& b) {
    const unsigned long long ia = (unsigned long long)__double_as_longlong(a);
    const unsigned long long ib = (unsigned long long)__double_as_longlong(b);
    if constexpr (S == 32) {
        const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)ia, (unsigned)ib, false, false);
        const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(ia >> 32), (unsigned)(ib >> 32), false, false);
        a = join_halves(lo[0], hi[0]);
        b = join_halves(lo[1], hi[1]);
    } else {
        const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)ia, (unsigned)ib, false, false);
        const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(ia >> 32), (unsigned)(ib >> 32), false, false);
        a = join_halves(lo[0], hi[0]);
        b = join_halves(lo[1], hi[1]);
    }
}
// Value of the partner lane for the exchange steps below 16, by DPP: row_ror:8 (lane ^ 8),
// row_half_mirror (lane ^ 7 within 8 lanes), quad_perm [2,3,0,1] (lane ^ 2), quad_perm [1,0,3,2] (lane ^ 1).
template <int S>
__device__ __forceinline__ double lane_partner(double v) {
    constexpr int ctrl = S == 8 ? 0x128 : S == 4 ? 0x141 : S == 2 ? 0x4E : 0xB1;
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), ctrl, 0xF, 0xF, false);
    return join_halves((unsigned)lo, (unsigned)hi);
}

// Sum of V[0..CNT) across the wavefront by recursive halving: at exchange step S a lane keeps the
// half of its values selected by bit S of its lane id and adds the partner's copy of that half.  The
// partners are lane ^ 32, ^ 16, ^ 8, ^ 7, ^ 2, ^ 1: the masks are independent, so the six steps cover
// all 64 lanes, and each partner differs from the lane in bit S.  Afterwards V[0] of lane L holds the
// total of value index L >> (6 - log2 CNT).  Steps 32 and 16 trade the two halves whole: after the
// swap the lane's kept half and the partner's copy of it are both in its registers (a + b).
template <int CNT, int S>
struct Halving {
    template <int NV>
    __device__ static __forceinline__ void run(double (&V)[NV], int lane) {
        if constexpr (CNT > 1) {
            constexpr int H = CNT / 2;
            if constexpr (S >= 16) {
#pragma unroll
                for (int t = 0; t < H; ++t) {
                    double a = V[t], b = V[H + t];
                    lane_swap<S>(a, b);
                    V[t] = a + b;
                }
            } else {
                const bool up = (lane & S) != 0;
#pragma unroll
                for (int t = 0; t < H; ++t) {
                    const double keep = up ? V[H + t] : V[t];
                    const double send = up ? V[t] : V[H + t];
                    V[t] = keep + lane_partner<S>(send);
                }
            }
            if constexpr (S > 1) Halving<H, S / 2>::run(V, lane);
        } else {
            if constexpr (S >= 16) {
                double a = V[0], b = V[0];
                lane_swap<S>(a, b);
                V[0] = a + b;
            } else {
                V[0] += lane_partner<S>(V[0]);
            }
            if constexpr (S > 1) Halving<1, S / 2>::run(V, lane);
        }
    }
};

template <int NV>
constexpr int log2c() {
    return NV <= 1 ? 0 : 1 + log2c<NV / 2>();
}

}  // namespace hfgpu
