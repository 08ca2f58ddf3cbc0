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
#include <type_traits>
#include <vector>

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

// Element loads of CPL consecutive columns: one aligned 16-byte (FP64) or 8-byte (FP32) load for a column pair,
// a plain load for one column.
template <typename T, int CPL>
struct ColsOf;
template <typename T>
struct ColsOf<T, 1> {
    using V = T;
    __device__ static double get(const V& v, int) { return (double)v; }
};
template <typename T>
struct ColsOf<T, 2> {
    using V = typename PairOf<T>::V;
    __device__ static double get(const V& v, int h) { return h ? (double)v.y : (double)v.x; }
};
template <int CPL>
struct DwOf;
template <>
struct DwOf<1> {
    using V = double;
    __device__ static double get(const V& v, int) { return v; }
};
template <>
struct DwOf<2> {
    using V = double2;
    __device__ static double get(const V& v, int h) { return h ? v.y : v.x; }
};

// Buffer-resource word 3 of gfx9-family raw buffers (32-bit data format, no swizzle, no index stride).
constexpr int kBufferFlags = 0x00020000;

// A raw buffer load of 4, 8 or 16 bytes at byte offset off; zero where off is beyond the buffer's range.
template <typename V>
__device__ __forceinline__ V buffer_load(__amdgpu_buffer_rsrc_t r, unsigned off) {
    if constexpr (sizeof(V) == 16)
        return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0));
    else if constexpr (sizeof(V) == 8)
        return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
    else
        return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}

// Blocks of NR store rows: bra pairs (i, j0 + s), s < NR, j0 = NR t, of one bra index i.  Row i holds
// ceil((i + 1) / NR) blocks, so blocks_before(i) = sum_{m=1..i} ceil(m / NR) = NR q (q + 1) / 2 + r (q + 1)
// with q = i / NR, r = i % NR.
__host__ __device__ inline int64_t blocks_before(int i, int nr) {
    const int64_t q = i / nr, r = i % nr;
    return nr * q * (q + 1) / 2 + r * (q + 1);
}

// Store rows the contraction can leave out: the store is screened (a shell quartet whose Schwarz product is below the
// threshold was never evaluated, so its elements are zero), so a row P is zero whenever qf_P qmax < thr, qf being an
// upper bound (rounded up to FP32) of the Schwarz factor of the position's shell pair, qmax the largest and thr the
// store's threshold less a margin.  Such a row is given an empty buffer range, whose loads return the same zeros
// without an HBM access, and a block of such rows returns at once, so the results are those of the full stream.
// qf null: every row.
struct StoreSkip {
    const float* qf;      // per padded pair position (0 for padding)
    double thr;
    double qmax;          // largest qf
};

// One workgroup per block of NR store rows P_s = pair_pad_index(i, j_s) with one bra index i.  A wavefront
// covers a row k of the pair triangle in chunks of W = 64 CPL columns: lane L owns columns CPL L + W m + h
// (h < CPL, m < MC = ceil(n / W)) and reads a chunk's CPL columns of each of the NR rows with one load each
// (two columns per lane halve the load instructions; one column per lane wastes fewer lanes on the short
// rows, row k having k + 1 elements).  The J weights Dw of a chunk are loaded once for the NR rows, and
// the rows share K_i: its row sums (sum_l B_s[k][l] D_{j_s}[l] over s) and its column part are one
// accumulator each, so a row k of the triangle needs NDENS (NR + 1) row sums instead of 2 NR NDENS.
// blk: the block's index among the blocks of this store (blocks_before(n, NR) of them).
template <int NDENS, int MC, int CPL, int ROWS, bool PREFETCH, int NR, typename T>
__device__ __forceinline__ void incore_jk_block(const T* __restrict__ G, int n, int64_t npad,
                                                const double* __restrict__ D, const double* __restrict__ Dw,
                                                double* __restrict__ J, double* __restrict__ K,
                                                long long* __restrict__ fxK, const StoreSkip& sk, int64_t blk) {
    extern __shared__ double lds[];
    constexpr int W = 64 * CPL;
    int i = (int)sqrt(2.0 * NR * (double)blk) - 1;
    if (i < 0) i = 0;
    while (blocks_before(i + 1, NR) <= blk) ++i;
    while (i > 0 && blocks_before(i, NR) > blk) --i;
    const int j0 = (int)(blk - blocks_before(i, NR)) * NR;
    const int64_t P0 = hfgpu::pair_pad_base(i) + j0;     // row of s = 0; rows s with j0 + s > i are empty
    const int nrow = i - j0 + 1 < NR ? i - j0 + 1 : NR;
    // store rows the screening zeroed entirely (qf_P qmax below the threshold) are not read; a block whose rows are
    // all zero has nothing to add
    bool live[NR];
    {
        bool any = false;
#pragma unroll
        for (int s = 0; s < NR; ++s) {
            live[s] = s < nrow && (!sk.qf || (double)sk.qf[P0 + s] * sk.qmax >= sk.thr);
            any = any || live[s];
        }
        if (!any) return;
    }
    double* xj = lds;                      // D_d[j_s][:]      [NR][NDENS][n]
    double* xi = xj + NR * NDENS * n;      // D_d[i][:]        [NDENS][n]
    double* rI = xi + NDENS * n;           // additions to K_d[i][:]
    double* rJ = rI + NDENS * n;           // additions to K_d[j_s][:]   [NR][NDENS][n]
    double* gdg = rJ + NR * NDENS * n;     // diagonal elements B_s[c][c]  [NR][n]
    double* red = gdg + NR * n;            // per-wavefront partials of J_{i j_s}  [NR][4]
    const size_t nn = (size_t)n * n;
    for (int t = threadIdx.x; t < NDENS * n; t += blockDim.x) {
        const int d = t / n, c = t - d * n;
        xi[t] = D[d * nn + (size_t)i * n + c];
#pragma unroll
        for (int s = 0; s < NR; ++s) xj[s * NDENS * n + t] = s < nrow ? D[d * nn + (size_t)(j0 + s) * n + c] : 0.0;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave-uniform: row bases in SGPRs
    double xjl[NR][NDENS][MC][CPL], xil[NDENS][MC][CPL], cI[NDENS][MC][CPL], cJ[NR][NDENS][MC][CPL];
#pragma unroll
    for (int d = 0; d < NDENS; ++d)
#pragma unroll
        for (int m = 0; m < MC; ++m)
#pragma unroll
            for (int h = 0; h < CPL; ++h) {
                const int l = CPL * lane + W * m + h;
                xil[d][m][h] = l < n ? xi[d * n + l] : 0.0;
                cI[d][m][h] = 0.0;
#pragma unroll
                for (int s = 0; s < NR; ++s) {
                    xjl[s][d][m][h] = l < n ? xj[(s * NDENS + d) * n + l] : 0.0;
                    cJ[s][d][m][h] = 0.0;
                }
            }
    constexpr int JN = NR == 1 ? 1 : NR == 2 ? 2 : 4;   // J partials, padded to a power of two
    static_assert(NR <= 4, "at most four store rows per block");
    double jacc[JN];
#pragma unroll
    for (int s = 0; s < JN; ++s) jacc[s] = 0.0;
    const char* __restrict__ dwb = (const char*)Dw;
    constexpr int NA = NR + 1;                     // row sums of one row k and density: K_i, then K_{j_s}
    constexpr int NV = ROWS * NA * NDENS;
    static_assert((NV & (NV - 1)) == 0, "row sums are reduced by recursive halving");
    constexpr int SH = 6 - log2c<NV>();
    using PV = typename ColsOf<T, CPL>::V;
    using DV = typename DwOf<CPL>::V;
    // The loads of a row group: every column chunk of its ROWS rows k0 + 4r in each of the NR store rows,
    // and of Dw, issued together, as buffer loads whose range is row k (its padded, even length; zero for
    // rows past the last one and for empty store rows).  Loads beyond the range return zero, so elements
    // outside the row need no masks, and no load leaves it.
    auto load = [&](int k0, PV (&g)[ROWS][NR][MC], DV (&dd)[ROWS][MC]) {
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
            const int k = k0 + 4 * r;
            const int kk = k < n ? k : n - 1;
            const int64_t b = hfgpu::pair_pad_base(kk);
            const int len = k < n ? (int)(hfgpu::pair_pad_base(kk + 1) - b) : 0;
            const auto rd = __builtin_amdgcn_make_buffer_rsrc((void*)(dwb + b * 8), (short)0, len * 8, kBufferFlags);
            // chunk m holds columns W m ..; row k has columns 0..k, so the chunks past it are zero and their loads
            // are not issued (k is the same for the whole wavefront)
#pragma unroll
            for (int m = 0; m < MC; ++m)
                dd[r][m] = m == 0 || W * m <= k ? buffer_load<DV>(rd, (unsigned)(CPL * lane + W * m) * 8u) : DV{};
#pragma unroll
            for (int s = 0; s < NR; ++s) {
                // an empty range for a row the store's screening zeroed: its loads return zero without a fetch
                const auto rg = __builtin_amdgcn_make_buffer_rsrc(
                    (void*)(G + (P0 + s) * npad + b), (short)0, s < nrow && live[s] ? len * (int)sizeof(T) : 0,
                    kBufferFlags);
#pragma unroll
                for (int m = 0; m < MC; ++m)
                    g[r][s][m] = m == 0 || W * m <= k
                                     ? buffer_load<PV>(rg, (unsigned)(CPL * lane + W * m) * (unsigned)sizeof(T))
                                     : PV{};
            }
        }
    };
    // The arithmetic and the row-sum reduction of the group of rows k0 + 4 r held in g, dd.
    auto group = [&](int k0, const PV (&g)[ROWS][NR][MC], const DV (&dd)[ROWS][MC]) {
        double V[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) V[v] = 0.0;
        double xkj[ROWS][NR][NDENS], xki[ROWS][NDENS];
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
            const int k = k0 + 4 * r;
            const int kk = k < n ? k : n - 1;
#pragma unroll
            for (int d = 0; d < NDENS; ++d) {
                xki[r][d] = xi[d * n + kk];
#pragma unroll
                for (int s = 0; s < NR; ++s) xkj[r][s][d] = xj[(s * NDENS + d) * n + kk];
            }
        }
        const int klast = k0 + 4 * (ROWS - 1) < n ? k0 + 4 * (ROWS - 1) : n - 1;
#pragma unroll
        for (int m = 0; m < MC; ++m) {
            if (W * m > klast) break;              // uniform across the wavefront
#pragma unroll
            for (int r = 0; r < ROWS; ++r)
#pragma unroll
                for (int h = 0; h < CPL; ++h) {
                    const double dw = DwOf<CPL>::get(dd[r][m], h);
#pragma unroll
                    for (int s = 0; s < NR; ++s) {
                        // zero outside row k; the diagonal element l = k enters both the row and the column
                        // part, and the epilogue takes one of the two out
                        const double e = ColsOf<T, CPL>::get(g[r][s][m], h);
                        jacc[s] = fma(e, dw, jacc[s]);
#pragma unroll
                        for (int d = 0; d < NDENS; ++d) {
                            V[(r * NA) * NDENS + d] = fma(e, xjl[s][d][m][h], V[(r * NA) * NDENS + d]);
                            V[(r * NA + 1 + s) * NDENS + d] = fma(e, xil[d][m][h], V[(r * NA + 1 + s) * NDENS + d]);
                            cI[d][m][h] = fma(e, xkj[r][s][d], cI[d][m][h]);
                            cJ[s][d][m][h] = fma(e, xki[r][d], cJ[s][d][m][h]);
                        }
                    }
                }
        }
        // the diagonal elements B_s[k][k] of this group's rows, for the epilogue: the lane that loaded column k
        // keeps it (the chunk and pair half are uniform across the wavefront)
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
            const int k = k0 + 4 * r;
            if (k < n && lane == (k % W) / CPL) {
                const int mk = k / W, hk = k % CPL;
#pragma unroll
                for (int s = 0; s < NR; ++s) {
                    double dv = 0.0;
#pragma unroll
                    for (int m = 0; m < MC; ++m)
#pragma unroll
                        for (int h = 0; h < CPL; ++h) {   // constant indices: selects, no indexed register access
                            const double x = ColsOf<T, CPL>::get(g[r][s][m], h);
                            dv = (m == mk && h == hk) ? x : dv;
                        }
                    gdg[s * n + k] = dv;
                }
            }
        }
        Halving<NV, 32>::run(V, lane);
        if ((lane & ((1 << SH) - 1)) == 0) {
            const int idx = lane >> SH;
            const int r = idx / (NA * NDENS), a = (idx / NDENS) % NA, d = idx % NDENS;
            const int k = k0 + 4 * r;
            // row k belongs to this wavefront alone
            if (k < n) (a == 0 ? rI + d * n : rJ + ((a - 1) * NDENS + d) * n)[k] = V[0];
        }
    };
    PV ga[ROWS][NR][MC];
    DV da[ROWS][MC];
    load(w, ga, da);
    if constexpr (PREFETCH) {
        // Two register buffers in turn: the next group's loads go out before this group's arithmetic and
        // reduction, into the buffer the previous group has finished with (no copies between buffers, so
        // nothing waits for the next group's data before its own arithmetic).
        PV gb[ROWS][NR][MC];
        DV db[ROWS][MC];
        for (int k0 = w; k0 < n;) {
            if (k0 + 4 * ROWS < n) load(k0 + 4 * ROWS, gb, db);
            group(k0, ga, da);
            k0 += 4 * ROWS;
            if (k0 >= n) break;
            if (k0 + 4 * ROWS < n) load(k0 + 4 * ROWS, ga, da);
            group(k0, gb, db);
            k0 += 4 * ROWS;
        }
    } else {
        for (int k0 = w; k0 < n; k0 += 4 * ROWS) {
            group(k0, ga, da);
            if (k0 + 4 * ROWS < n) load(k0 + 4 * ROWS, ga, da);
        }
    }
    // J_{i j_s}: wavefront sums, then the block's
    Halving<JN, 32>::run(jacc, lane);               // lane L then holds the sum of jacc[L >> (6 - log2 JN)]
    {
        constexpr int JS = 6 - log2c<JN>();
        if ((lane & ((1 << JS) - 1)) == 0 && (lane >> JS) < NR) red[(lane >> JS) * 4 + w] = jacc[0];
    }
    __syncthreads();
    // column parts of every wavefront (deterministic mode: the wavefronts in turn, so that every element
    // receives their parts in one order)
    for (int turn = 0; turn < (fxK ? 4 : 1); ++turn) {
        if (!fxK || w == turn) {
#pragma unroll
            for (int d = 0; d < NDENS; ++d)
#pragma unroll
                for (int m = 0; m < MC; ++m)
#pragma unroll
                    for (int h = 0; h < CPL; ++h) {
                        const int l = CPL * lane + W * m + h;
                        if (l < n) {
                            atomicAdd(&rI[d * n + l], cI[d][m][h]);
#pragma unroll
                            for (int s = 0; s < NR; ++s) atomicAdd(&rJ[(s * NDENS + d) * n + l], cJ[s][d][m][h]);
                        }
                    }
        }
        __syncthreads();
    }
    if (threadIdx.x < nrow) {
        const int s = threadIdx.x, j = j0 + s;
        const double v = (red[s * 4] + red[s * 4 + 1]) + (red[s * 4 + 2] + red[s * 4 + 3]);
        J[(size_t)i * n + j] += v;
        if (i != j) J[(size_t)j * n + i] += v;
    }
    // the diagonal element B_s[c][c] was added to K row c by both parts: one of them comes out here
    for (int t = threadIdx.x; t < NDENS * n; t += blockDim.x) {
        const int d = t / n, c = t - d * n;
        double vi = rI[t];
#pragma unroll
        for (int s = 0; s < NR; ++s) {
            if (s >= nrow) break;
            const int j = j0 + s;
            vi = fma(-gdg[s * n + c], xj[s * NDENS * n + t], vi);
            if (i != j) {
                const double vj = fma(-gdg[s * n + c], xi[t], rJ[s * NDENS * n + t]);
                if (fxK)
                    hfgpu::fx_add(fxK, d * nn + (size_t)j * n + c, vj);
                else
                    atomicAdd(&K[d * nn + (size_t)j * n + c], vj);
            }
        }
        if (fxK)
            hfgpu::fx_add(fxK, d * nn + (size_t)i * n + c, vi);
        else
            atomicAdd(&K[d * nn + (size_t)i * n + c], vi);
    }
}

template <int NDENS, int MC, int CPL, int ROWS, bool PREFETCH, int NR, typename T>
__global__ void __launch_bounds__(256) incore_jk_kernel(const T* __restrict__ G, int n, int64_t npad,
                                                        const double* __restrict__ D, const double* __restrict__ Dw,
                                                        double* __restrict__ J, double* __restrict__ K,
                                                        long long* __restrict__ fxK, StoreSkip sk) {
    incore_jk_block<NDENS, MC, CPL, ROWS, PREFETCH, NR, T>(G, n, npad, D, Dw, J, K, fxK, sk, blockIdx.x);
}

// The contractions of several stores in one launch (the molecules of a lockstep batch): job j owns blocks
// [first[j], first[j + 1]) of the grid, each store with its own density, weights, outputs and row skips.  The
// jobs of one launch share the kernel's shape (columns per lane, chunks, rows per group, store rows per block);
// the dynamic LDS is sized for the largest n.
constexpr int kIncoreBatch = 32;
template <typename T>
struct IncoreBatchArgs {
    int count;
    int64_t first[kIncoreBatch + 1];
    int n[kIncoreBatch];
    const T* G[kIncoreBatch];
    const double* D[kIncoreBatch];
    const double* Dw[kIncoreBatch];
    double* J[kIncoreBatch];
    double* K[kIncoreBatch];
    StoreSkip sk[kIncoreBatch];
};

template <int NDENS, int MC, int CPL, int ROWS, bool PREFETCH, int NR, typename T>
__global__ void __launch_bounds__(256) incore_jk_batch_kernel(IncoreBatchArgs<T> g) {
    const int64_t b = blockIdx.x;
    int j = 0;
    while (j + 1 < g.count && b >= g.first[j + 1]) ++j;
    const int n = g.n[j];
    incore_jk_block<NDENS, MC, CPL, ROWS, PREFETCH, NR, T>(g.G[j], n, hfgpu::pair_pad_base(n), g.D[j], g.Dw[j], g.J[j],
                                                           g.K[j], nullptr, g.sk[j], b - g.first[j]);
}

// Deterministic mode: K (+)= the fixed-point sums
__global__ void incore_fx_finalize_kernel(const long long* __restrict__ fx, double* __restrict__ K, size_t len,
                                          int accumulate) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (size_t)gridDim.x * blockDim.x) {
        const double v = hfgpu::fx_value(fx[2 * i], fx[2 * i + 1]);
        K[i] = accumulate ? K[i] + v : v;
    }
}

// Launches of the single-store kernel and of the batched one: run<...>() launches the instance for MC column
// chunks.  Only instances that the in-core size limit (n <= kIncoreMaxN) and the register bounds of shape_fits can
// reach are compiled.
template <int NDENS, int MC, int CPL, int NR>
constexpr bool kIncoreInstance = 64 * CPL * (MC - 1) < kIncoreMaxN && (CPL == 2 || NDENS == 1 || MC <= 4) &&
                                 (NR == 1 || NDENS == 1 || 64 * CPL * (MC - 1) < 256);

inline size_t incore_lds_bytes(int ndens, int nr, int n) {
    return ((2 * nr + 2) * (size_t)ndens * n + (size_t)nr * n + 4 * nr) * sizeof(double);
}

template <typename T>
struct OneStore {
    const T* G;
    int n;
    int64_t npad;
    const double* D;
    const double* Dw;
    double* J;
    double* K;
    long long* fxK;
    hipStream_t st;
    StoreSkip sk;
    template <int NDENS, int MC, int CPL, int ROWS, bool PREFETCH, int NR>
    hipError_t run() const {
        if constexpr (kIncoreInstance<NDENS, MC, CPL, NR>) {
            hipLaunchKernelGGL((incore_jk_kernel<NDENS, MC, CPL, ROWS, PREFETCH, NR, T>),
                               dim3((unsigned)blocks_before(n, NR)), dim3(256), incore_lds_bytes(NDENS, NR, n), st, G, n,
                               npad, D, Dw, J, K, fxK, sk);
            return hipGetLastError();
        } else {
            return hipErrorInvalidValue;
        }
    }
};

// g->first[] is filled by run() for the shape's NR.
template <typename T>
struct ManyStores {
    IncoreBatchArgs<T>* g;
    int nmax;
    hipStream_t st;
    template <int NDENS, int MC, int CPL, int ROWS, bool PREFETCH, int NR>
    hipError_t run() const {
        if constexpr (kIncoreInstance<NDENS, MC, CPL, NR>) {
            int64_t b = 0;
            for (int j = 0; j < g->count; ++j) {
                g->first[j] = b;
                b += blocks_before(g->n[j], NR);
            }
            g->first[g->count] = b;
            hipLaunchKernelGGL((incore_jk_batch_kernel<NDENS, MC, CPL, ROWS, PREFETCH, NR, T>), dim3((unsigned)b),
                               dim3(256), incore_lds_bytes(NDENS, NR, nmax), st, *g);
            return hipGetLastError();
        } else {
            return hipErrorInvalidValue;
        }
    }
};

template <int NDENS, int CPL, int ROWS, bool PREFETCH, int NR, class Go>
hipError_t launch_incore(int n, const Go& go) {
    switch ((n + 64 * CPL - 1) / (64 * CPL)) {
        case 1: return go.template run<NDENS, 1, CPL, ROWS, PREFETCH, NR>();
        case 2: return go.template run<NDENS, 2, CPL, ROWS, PREFETCH, NR>();
        case 3: return go.template run<NDENS, 3, CPL, ROWS, PREFETCH, NR>();
        case 4: return go.template run<NDENS, 4, CPL, ROWS, PREFETCH, NR>();
        case 5: return go.template run<NDENS, 5, CPL, ROWS, PREFETCH, NR>();
        case 6: return go.template run<NDENS, 6, CPL, ROWS, PREFETCH, NR>();
        case 7: return go.template run<NDENS, 7, CPL, ROWS, PREFETCH, NR>();
        case 8: return go.template run<NDENS, 8, CPL, ROWS, PREFETCH, NR>();
        default: return hipErrorInvalidValue;
    }
}

// Tile shapes of the contraction: columns per lane, rows k of the pair triangle whose loads are issued
// together and whose row sums are reduced together, whether the next group's loads are issued before the
// current group's arithmetic (two register buffers in turn: more bytes in flight per wavefront for more
// registers, so fewer wavefronts), and store rows per block (NR).
struct ShapeDesc {
    int cpl, rows, prefetch, nr;
};
constexpr ShapeDesc kShapes[] = {
    {2, 2, 0, 1}, {2, 2, 1, 1}, {2, 1, 1, 1}, {1, 2, 0, 1}, {1, 2, 1, 1}, {2, 1, 0, 3}, {2, 1, 1, 3}, {2, 2, 0, 3},
};
constexpr int kIncoreShapes = (int)(sizeof(kShapes) / sizeof(kShapes[0]));

// Whether a shape has an instance for n and the density count: one column per lane needs ceil(n / 64)
// column registers per density and array (at most 8, n <= 512, with one density; 4, n <= 256, with two),
// and three store rows per block with two densities are kept to n <= 256 for registers.
inline bool shape_fits(int shape, int n, int ndens) {
    const ShapeDesc& d = kShapes[shape];
    if (d.cpl == 1 && n > (ndens == 1 ? 512 : 256)) return false;
    if (d.nr > 1 && ndens == 2 && n > 256) return false;
    return true;
}

// The shape a launch for n functions and ndens densities uses (the default where the asked one does not fit).
inline int shape_for(int shape, int n, int ndens) { return shape_fits(shape, n, ndens) ? shape : 0; }

// Column chunks of the kernel instance that shape runs with for n functions.
inline int shape_chunks(int shape, int n) {
    const int w = 64 * kShapes[shape].cpl;
    return (n + w - 1) / w;
}

// n: the size that selects the column chunks (the largest store of a batched launch).
template <int NDENS, class Go>
hipError_t launch_incore_nd(int shape, int n, const Go& go) {
    switch (shape_for(shape, n, NDENS)) {
        case 1: return launch_incore<NDENS, 2, 2, true, 1>(n, go);
        case 2: return launch_incore<NDENS, 2, 1, true, 1>(n, go);
        case 3: return launch_incore<NDENS, 1, 2, false, 1>(n, go);
        case 4: return launch_incore<NDENS, 1, 2, true, 1>(n, go);
        case 5: return launch_incore<NDENS, 2, 1, false, 3>(n, go);
        case 6: return launch_incore<NDENS, 2, 1, true, 3>(n, go);
        case 7: return launch_incore<NDENS, 2, 2, false, 3>(n, go);
        default: return launch_incore<NDENS, 2, 2, false, 1>(n, go);
    }
}

template <typename T>
hipError_t launch_incore_shape(int shape, int ndens, const T* G, int n, int64_t np, const double* D,
                               const double* Dw, double* J, double* K, long long* fxK, hipStream_t st,
                               const StoreSkip& sk) {
    const OneStore<T> go{G, n, np, D, Dw, J, K, fxK, st, sk};
    return ndens == 1 ? launch_incore_nd<1>(shape, n, go) : launch_incore_nd<2>(shape, n, go);
}

template <typename T>
hipError_t launch_incore_batch(int shape, int ndens, IncoreBatchArgs<T>& g, int nmax, hipStream_t st) {
    const ManyStores<T> go{&g, nmax, st};
    return ndens == 1 ? launch_incore_nd<1>(shape, nmax, go) : launch_incore_nd<2>(shape, nmax, go);
}

// Autotuner: the tile shape is picked per (64-column chunks ceil(n/64), densities, store
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

int tuning_key(int n, int ndens, bool fp32) { return ((n + 63) / 64) * 100 + ndens * 10 + (fp32 ? 1 : 0); }

// The shape the autotuner picked for this kind of build, or the default when it has not run (no timing here).
int known_shape(int n, int ndens, bool fp32) {
    std::lock_guard<std::mutex> lock(tuning().mu);
    auto it = tuning().shape.find(tuning_key(n, ndens, fp32));
    return it == tuning().shape.end() ? 0 : it->second;
}

template <typename T>
int tuned_shape(hf_ctx*, int ndens, const T* G, int n, int64_t np, const double* D, const double* Dw,
               hipStream_t st, const StoreSkip& sk) {
    const int key = tuning_key(n, ndens, sizeof(T) == 4);
    {
        std::lock_guard<std::mutex> lock(tuning().mu);
        auto it = tuning().shape.find(key);
        if (it != tuning().shape.end()) return it->second;
    }
    const size_t nn = (size_t)n * n;
    DevBuf<double> sJ, sK;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int best = 0;
    float best_ms = 0;
    bool ok = sJ.alloc(nn) == hipSuccess && sK.alloc(2 * nn) == hipSuccess &&
              hipEventCreate(&ev[0]) == hipSuccess && hipEventCreate(&ev[1]) == hipSuccess;
    for (int shape = 0; shape < kIncoreShapes; ++shape) {
        if (!ok) break;
        if (!shape_fits(shape, n, ndens)) continue;
        float t = 0;
        for (int rep = 0; rep < 3 && ok; ++rep) {          // first run warms the caches and the code
            ok = hipEventRecord(ev[0], st) == hipSuccess &&
                 launch_incore_shape(shape, ndens, G, n, np, D, Dw, sJ.p, sK.p, nullptr, st, sk) == hipSuccess &&
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
    if (!ok) return 0;                                     // tuning failed: the default, not recorded
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

// Kernel statistics of the contraction of one store: the elements streamed (the rows the screening left, all rows
// without skipping) and their flops; launch: the launch is counted here too (a batched launch counts once).
void incore_account(hf_ctx* ctx, int f, const StoreSkip& sk, int64_t np, int ndens, bool launch) {
    KernelStat& ks = ctx->incore_stat[f];
    if (launch) ks.launches += 1;
    const double rows = sk.qf ? (double)ctx->qf_live : (double)np;
    ks.eris += rows * (double)np;
    ks.flops += rows * (double)np * (2.0 + 8.0 * ndens);                 // J, and row + column parts of K
}

// profile = 1: HIP events around a contraction launch on st; its time is read at the next launch's begin (or the
// context's next statistics flush) and added to incore_stat[f].
int incore_events_begin(hf_ctx* ctx, hipStream_t st) {
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
    return HF_OK;
}

int incore_events_end(hf_ctx* ctx, hipStream_t st, int f) {
    HF_HIP_CHECK(ctx, hipEventRecord(ctx->incore_ev[1], st));
    ctx->incore_pending = f;
    return HF_OK;
}

// The contraction of one in-core J/K build, after its packed weighted density is in Dw (np doubles) and, unless
// accumulate, J and K are zero.  no_alloc: the tile shape is the one already tuned (or the default), and no events
// are created or recorded, so nothing is allocated.
int incore_contract(hf_ctx* ctx, int ndens, const double* G, const float* g32, const double* d_D, double* d_J,
                    double* d_K, hipStream_t st, bool accumulate, double* Dw, bool no_alloc, const StoreSkip& sk) {
    const int n = ctx->mol.nbf;
    const int64_t np = hfgpu::pair_pad_base(n);      // padded pair positions (rows and columns of the store)
    const size_t nn = (size_t)n * n;
    // tile shape from the autotuner (first build of each kind in the process times the candidates)
    const int forced = ctx->opt.incore_shape;
    if (forced >= kIncoreShapes) return ctx->fail(HF_E_ARG, "incore_shape must be -1 or 0.." +
                                                             std::to_string(kIncoreShapes - 1));
    const int shape = forced >= 0 ? forced
                      : no_alloc  ? known_shape(n, ndens, g32 != nullptr)
                      : g32    ? tuned_shape(ctx, ndens, g32, n, np, d_D, Dw, st, sk)
                               : tuned_shape(ctx, ndens, G, n, np, d_D, Dw, st, sk);
    // statistics, and with profile = 1 events around the contraction (read at the next flush)
    const int f = g32 ? 1 : 0;
    const bool prof = ctx->opt.profile != 0 && !no_alloc;
    if (prof)
        if (int rc = incore_events_begin(ctx, st)) return rc;
    incore_account(ctx, f, sk, np, ndens, true);
    // deterministic mode: K accumulates in fixed point (fx_add) and is folded in after the contraction
    long long* fxK = nullptr;
    if (ctx->opt.deterministic) {
        if (no_alloc)
            return ctx->fail(HF_E_UNSUPPORTED, "deterministic mode needs an accumulator the allocation-free entry "
                                               "point does not have");
        HF_HIP_CHECK(ctx, ctx->d_fx.alloc(2 * 3 * nn));
        hipLaunchKernelGGL(incore_fill_kernel, dim3((unsigned)((2 * ndens * nn + 255) / 256)), dim3(256), 0, st,
                           (double*)ctx->d_fx.p, 2 * ndens * nn);
        fxK = ctx->d_fx.p;
    }
    hipError_t e;
    if (g32)
        e = launch_incore_shape(shape, ndens, g32, n, np, d_D, Dw, d_J, d_K, fxK, st, sk);
    else
        e = launch_incore_shape(shape, ndens, G, n, np, d_D, Dw, d_J, d_K, fxK, st, sk);
    if (e != hipSuccess) return ctx->fail(HF_E_HIP, std::string("in-core J/K: ") + hipGetErrorString(e));
    if (fxK)
        hipLaunchKernelGGL(incore_fx_finalize_kernel, dim3((unsigned)std::min<size_t>((ndens * nn + 255) / 256, 4096)),
                           dim3(256), 0, st, fxK, d_K, ndens * nn, accumulate ? 1 : 0);
    if (prof)
        if (int rc = incore_events_end(ctx, st, f)) return rc;
    return HF_OK;
}


// One in-core J/K build: the packed weighted density, zero outputs (unless accumulate), the contraction.
int incore_jk_run(hf_ctx* ctx, int ndens, const double* G, const float* g32, const double* d_D, double* d_J,
                  double* d_K, hipStream_t st, bool accumulate, double* Dw, bool no_alloc, const StoreSkip& sk) {
    const int n = ctx->mol.nbf;
    const size_t nn = (size_t)n * n;
    hipLaunchKernelGGL(incore_dw_kernel, dim3(n), dim3(256), 0, st, d_D, n, ndens, Dw);
    if (!accumulate) {
        hipLaunchKernelGGL(incore_fill_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, d_J, nn);
        hipLaunchKernelGGL(incore_fill_kernel, dim3((unsigned)((ndens * nn + 255) / 256)), dim3(256), 0, st, d_K,
                           ndens * nn);
    }
    return incore_contract(ctx, ndens, G, g32, d_D, d_J, d_K, st, accumulate, Dw, no_alloc, sk);
}

// The skips of a store this context evaluated (at its Schwarz threshold), or none.
StoreSkip store_skip(const hf_ctx* ctx, bool own_store) {
    StoreSkip sk{nullptr, 0.0, 0.0};
    if (own_store && ctx->qf_valid && ctx->opt.schwarz_threshold > 0) {
        sk.qf = ctx->d_qf.p;
        // the task lists keep a quartet when Q_ket >= thr / Q_bra; the margin covers the rounding of that division
        sk.thr = ctx->opt.schwarz_threshold * (1.0 - 1e-6);
        sk.qmax = ctx->qf_max;
    }
    return sk;
}

// The set-up launches of several in-core builds in one: per job (blockIdx.y) the packed weighted density and,
// unless accumulate, zero J and K (the same values incore_dw_kernel and incore_fill_kernel write).
constexpr int kPrepJobs = 32;
struct IncorePrep {
    int count;
    int n[kPrepJobs], ndens[kPrepJobs], clear[kPrepJobs];
    const double* D[kPrepJobs];
    double* Dw[kPrepJobs];
    double* J[kPrepJobs];
    double* K[kPrepJobs];
};

__global__ void __launch_bounds__(256) incore_prep_batch_kernel(IncorePrep g) {
    const int j = blockIdx.y, n = g.n[j], ndens = g.ndens[j];
    const size_t nn = (size_t)n * n;
    const double* __restrict__ D = g.D[j];
    double* __restrict__ Dw = g.Dw[j];
    for (int k = blockIdx.x; k < n; k += gridDim.x) {
        const int64_t b = hfgpu::pair_pad_base(k), len = hfgpu::pair_pad_base(k + 1) - b;
        for (int l = threadIdx.x; l < len; l += blockDim.x) {
            double v = 0.0;
            if (l <= k)
                for (int d = 0; d < ndens; ++d) v += D[d * nn + (size_t)k * n + l];
            Dw[b + l] = (l == k ? 1.0 : 2.0) * v;
        }
    }
    if (g.clear[j]) {
        double* __restrict__ J = g.J[j];
        double* __restrict__ K = g.K[j];
        for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nn; idx += (size_t)gridDim.x * blockDim.x)
            J[idx] = 0.0;
        for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < ndens * nn;
             idx += (size_t)gridDim.x * blockDim.x)
            K[idx] = 0.0;
    }
}

}  // namespace

int hf_incore_jk_impl(hf_ctx* ctx, int ndens, const double* G, const float* g32, const double* d_D, double* d_J,
                      double* d_K, hipStream_t st, bool accumulate, bool own_store) {
    if (!ctx->have_mol) return ctx->fail(HF_E_STATE, "no molecule set");
    if (ndens < 1 || ndens > 2) return ctx->fail(HF_E_ARG, "ndens must be 1 or 2");
    const int n = ctx->mol.nbf;
    if (n > kIncoreMaxN) return ctx->fail(HF_E_UNSUPPORTED, "in-core J/K takes at most 512 basis functions");
    HF_HIP_CHECK(ctx, ctx->d_Dw.alloc((size_t)hfgpu::pair_pad_base(n)));
    return incore_jk_run(ctx, ndens, G, g32, d_D, d_J, d_K, st, accumulate, ctx->d_Dw.p, false,
                         store_skip(ctx, own_store));
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
                         Dw, true, store_skip(ctx, false));
}

namespace {

// One launch of the batched contraction over jobs[grp] (same store precision T, densities and tile shape, at most
// kIncoreBatch jobs), after the set-up launch of hf_incore_jk_batch.  Statistics go to each job's context; the
// launch (and, with profile = 1, its time) to the first job's.
template <typename T>
int incore_contract_many(const std::vector<IncoreJob>& jobs, const std::vector<size_t>& grp, int shape,
                         hipStream_t st) {
    hf_ctx* lead = jobs[grp[0]].ctx;
    IncoreBatchArgs<T> g{};
    g.count = (int)grp.size();
    int nmax = 1;
    const int ndens = jobs[grp[0]].ndens;
    const int f = std::is_same<T, float>::value ? 1 : 0;
    for (int j = 0; j < g.count; ++j) {
        const IncoreJob& jb = jobs[grp[j]];
        hf_ctx* ctx = jb.ctx;
        const int n = ctx->mol.nbf;
        g.n[j] = n;
        if constexpr (std::is_same<T, float>::value)
            g.G[j] = jb.g32;
        else
            g.G[j] = jb.G;
        g.D[j] = jb.D;
        g.Dw[j] = ctx->d_Dw.p;
        g.J[j] = jb.J;
        g.K[j] = jb.K;
        g.sk[j] = store_skip(ctx, jb.own_store);
        nmax = std::max(nmax, n);
        incore_account(ctx, f, g.sk[j], hfgpu::pair_pad_base(n), ndens, j == 0);
    }
    const bool prof = lead->opt.profile != 0;
    if (prof)
        if (int rc = incore_events_begin(lead, st)) return rc;
    const hipError_t e = launch_incore_batch(shape, ndens, g, nmax, st);
    if (e != hipSuccess) return lead->fail(HF_E_HIP, std::string("batched in-core J/K: ") + hipGetErrorString(e));
    if (prof)
        if (int rc = incore_events_end(lead, st, f)) return rc;
    return HF_OK;
}

}  // namespace

int hf_incore_jk_batch(const std::vector<IncoreJob>& jobs, hipStream_t st) {
    if (jobs.empty()) return HF_OK;
    for (const IncoreJob& q : jobs) {
        hf_ctx* ctx = q.ctx;
        if (!ctx->have_mol) return ctx->fail(HF_E_STATE, "no molecule set");
        if (q.ndens < 1 || q.ndens > 2) return ctx->fail(HF_E_ARG, "ndens must be 1 or 2");
        if (ctx->mol.nbf > kIncoreMaxN) return ctx->fail(HF_E_UNSUPPORTED, "in-core J/K takes at most 512 basis functions");
        HF_HIP_CHECK(ctx, ctx->d_Dw.alloc((size_t)hfgpu::pair_pad_base(ctx->mol.nbf)));
    }
    for (size_t j0 = 0; j0 < jobs.size(); j0 += kPrepJobs) {
        IncorePrep g{};
        g.count = (int)std::min<size_t>(kPrepJobs, jobs.size() - j0);
        int nmax = 1;
        for (int j = 0; j < g.count; ++j) {
            const IncoreJob& q = jobs[j0 + j];
            g.n[j] = q.ctx->mol.nbf;
            g.ndens[j] = q.ndens;
            g.clear[j] = q.accumulate ? 0 : 1;
            g.D[j] = q.D;
            g.Dw[j] = q.ctx->d_Dw.p;
            g.J[j] = q.J;
            g.K[j] = q.K;
            nmax = std::max(nmax, g.n[j]);
        }
        hipLaunchKernelGGL(incore_prep_batch_kernel, dim3(nmax, g.count), dim3(256), 0, st, g);
    }
    HF_HIP_CHECK(jobs[0].ctx, hipGetLastError());
    // Contractions.  Jobs of one kind — store precision, densities, tile shape and column chunks — run as one
    // launch of the batched kernel; a job alone in its kind, and every job in deterministic mode (whose K goes
    // to its own fixed-point accumulator), runs its own launch.
    struct Item {
        size_t q;
        int key, shape;
    };
    std::vector<Item> items;
    std::vector<size_t> single;
    for (size_t q = 0; q < jobs.size(); ++q) {
        const IncoreJob& jb = jobs[q];
        hf_ctx* ctx = jb.ctx;
        const int forced = ctx->opt.incore_shape;
        if (ctx->opt.deterministic || forced >= kIncoreShapes) {
            single.push_back(q);
            continue;
        }
        const int n = ctx->mol.nbf;
        const int64_t np = hfgpu::pair_pad_base(n);
        const StoreSkip sk = store_skip(ctx, jb.own_store);
        int shape = forced >= 0 ? forced
                    : jb.g32    ? tuned_shape(ctx, jb.ndens, jb.g32, n, np, jb.D, ctx->d_Dw.p, st, sk)
                                : tuned_shape(ctx, jb.ndens, jb.G, n, np, jb.D, ctx->d_Dw.p, st, sk);
        shape = shape_for(shape, n, jb.ndens);
        const int key = (((jb.g32 ? 1 : 0) * 2 + jb.ndens - 1) * 16 + shape) * 16 + shape_chunks(shape, n);
        items.push_back({q, key, shape});
    }
    std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) { return a.key < b.key; });
    for (size_t i0 = 0; i0 < items.size();) {
        size_t i1 = i0 + 1;
        while (i1 < items.size() && items[i1].key == items[i0].key) ++i1;
        if (i1 - i0 == 1) {
            single.push_back(items[i0].q);
        } else {
            for (size_t c0 = i0; c0 < i1; c0 += kIncoreBatch) {
                const size_t c1 = std::min(i1, c0 + kIncoreBatch);
                std::vector<size_t> grp;
                for (size_t i = c0; i < c1; ++i) grp.push_back(items[i].q);
                const int rc = jobs[grp[0]].g32 ? incore_contract_many<float>(jobs, grp, items[c0].shape, st)
                                                : incore_contract_many<double>(jobs, grp, items[c0].shape, st);
                if (rc) return rc;
            }
        }
        i0 = i1;
    }
    for (size_t q : single) {
        const IncoreJob& jb = jobs[q];
        if (int rc = incore_contract(jb.ctx, jb.ndens, jb.G, jb.g32, jb.D, jb.J, jb.K, st, jb.accumulate,
                                     jb.ctx->d_Dw.p, false, store_skip(jb.ctx, jb.own_store)))
            return rc;
    }
    return HF_OK;
}
