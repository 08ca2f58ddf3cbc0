// SCF steps of several molecules in lockstep: every phase of an iteration — Fock assembly, the DIIS
// products and inner products, the orthogonalisation, SP2 purification, the density products, the energy
// reductions and the density update — is one launch over all the molecules (a "job" is one molecule and one
// spin), so a batch of B molecules takes the launches and host round trips of one.  The per-molecule J/K
// builds (in-core contraction or direct class kernels) are launched one molecule after another on the same
// stream.  The one-molecule SCF (hf_scf, hf_scf_step) runs these phases over a batch of one, so a molecule's
// numbers do not depend on the batch it ran in (up to the FP64 atomics' order in its J/K builds).
//
// The step follows RestrictedHartreeFock::Step / UnrestrictedHartreeFock::Step
// (HartreeFock/RestrictedHartreeFock.cpp:61-166, HartreeFock/UnrestrictedHartreeFock.cpp:87-191) and
// DIISStep (HartreeFock/RestrictedHartreeFock.cpp:91-120, DIIS.h:9-96); the density comes from SP2
// purification (the same aufbau projector as the reference's diagonalisation), diagonalisation being the
// fallback for a step it cannot resolve.
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>

#include "dense_tile.h"
#include "gershgorin.h"
#include "scf_run.h"
#include "wave_reduce.h"

using hfgpu::v4d;

namespace {

constexpr int kJ = kBatchJobs;
constexpr int kDotBlocks = 64;                              // blocks per reduction (fixed: results do not
                                                            // depend on the batch)
constexpr int kHistStride = 8 + 2 * (kSp2MaxSteps + 1);     // per SP2 job: lo, hi, final, then e_s, stop_s
constexpr int kSp2MaxTiles = (kSmallDense + 15) / 16;       // row tiles of the largest batched matrix
constexpr int kPartStride = 2 * kSp2MaxTiles * (kSp2MaxSteps + 1);

// ---------------------------------------------------------------- products
// C_j = alpha_j A_j B_j for square n_j x n_j column-major matrices; block b of the launch is tile
// b - tile_off[j] of job j (tiles in the order of dgemm16: row tile fastest).
struct BGemm {
    int count;
    int tile_off[kJ + 1];
    int n[kJ];
    double alpha[kJ];
    const double* A[kJ];
    const double* B[kJ];
    double* C[kJ];
};

__device__ __forceinline__ int job_of_block(const int* tile_off, int count) {
    int j = 0;
    while (j + 1 < count && (int)blockIdx.x >= tile_off[j + 1]) ++j;
    return j;
}

__global__ void __launch_bounds__(256) bgemm16_kernel(BGemm g) {
    __shared__ v4d part[3][64];
    const int j = job_of_block(g.tile_off, g.count);
    const int n = g.n[j], nt = (n + 15) / 16, t = (int)blockIdx.x - g.tile_off[j];
    hfgpu::gemm_tile<false, false>(n, n, n, g.alpha[j], g.A[j], n, g.B[j], n, 0.0, g.C[j], n, (t % nt) * 16,
                                   (t / nt) * 16, part);
}

// ---------------------------------------------------------------- element-wise
enum ElemOp { kOpFock, kOpAntisymCopy, kOpSym, kOpMix, kOpZero, kOpSub, kOpCopy };

struct BElem {
    int count, op;
    int n[kJ];
    size_t cnt[kJ];            // elements (n^2, or nsp n^2 for the density update)
    double x[kJ];              // kOpFock: K scale; kOpMix: alpha
    int flag[kJ];              // kOpFock: symmetrise; kOpMix: replace
    const double* a[kJ];
    const double* b[kJ];
    const double* c[kJ];
    double* out[kJ];
    double* out2[kJ];
};

__global__ void __launch_bounds__(256) belem_kernel(BElem g) {
    const int j = blockIdx.y, n = g.n[j];
    const size_t cnt = g.cnt[j];
    const double* __restrict__ a = g.a[j];
    const double* __restrict__ b = g.b[j];
    const double* __restrict__ c = g.c[j];
    double* __restrict__ out = g.out[j];
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < cnt; idx += (size_t)gridDim.x * blockDim.x) {
        switch (g.op) {
            case kOpFock: {   // F = h + J - x K, symmetrised (RHF) as the reference's Fock matrix
                const size_t i = idx / n, jj = idx % n, t = jj * n + i;
                const double v = a[idx] + b[idx] - g.x[j] * c[idx];
                out[idx] = g.flag[j] ? 0.5 * (v + (a[t] + b[t] - g.x[j] * c[t])) : v;
                break;
            }
            case kOpAntisymCopy: {   // DIIS error e = Y - Y^T (column-major Y = S D F), and F into its slot
                const size_t i = idx % n, jj = idx / n;
                out[idx] = a[idx] - a[i * n + jj];
                g.out2[j][idx] = b[idx];
                break;
            }
            case kOpSym: {
                const size_t i = idx % n, jj = idx / n;
                if (i < jj) {
                    const double v = 0.5 * (out[idx] + out[i * n + jj]);
                    out[idx] = v;
                    out[i * n + jj] = v;
                }
                break;
            }
            case kOpMix:
                out[idx] = g.flag[j] ? a[idx] : g.x[j] * a[idx] + (1.0 - g.x[j]) * out[idx];
                break;
            case kOpZero: out[idx] = 0.0; break;
            case kOpSub: out[idx] = a[idx] - b[idx]; break;
            default: out[idx] = a[idx]; break;
        }
    }
}

// DIIS extrapolation F = sum_r c_r F_slot(r) (HartreeFock/DIIS.h:70-94), slots oldest first
struct BLin {
    int count;
    size_t nn[kJ];
    const double* base[kJ];
    double* out[kJ];
    int m[kJ];
    unsigned char slot[kJ][8];
    double c[kJ][Diis::kKeep];
};

__global__ void __launch_bounds__(256) blin_kernel(BLin g) {
    const int j = blockIdx.y, m = g.m[j];
    const size_t nn = g.nn[j];
    const double* __restrict__ base = g.base[j];
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nn; idx += (size_t)gridDim.x * blockDim.x) {
        double v = 0;
        for (int r = 0; r < m; ++r) v += g.c[j][r] * base[(size_t)g.slot[j][r] * nn + idx];
        g.out[j][idx] = v;
    }
}

// ---------------------------------------------------------------- reductions
// Per job up to 6 reductions of two n x n matrices: x . y_r (kind 0), sum (x - y_r)^2 (1), max |x - y_r| (2).
// kDotBlocks blocks split each job's elements into contiguous chunks; a final wavefront per job combines
// the partials in a fixed order, so the sums are deterministic and independent of the batch.
struct BRed {
    int count, job0;           // results go to out[(job0 + j) * 8 + r]
    size_t nn[kJ];
    const double* x[kJ];
    const double* y[kJ][6];
    unsigned char kind[kJ][6];
    int m[kJ];
};

__global__ void __launch_bounds__(256) bred_partial_kernel(BRed g, double* __restrict__ part) {
    __shared__ double sh[256];
    const int j = blockIdx.y;
    const size_t nn = g.nn[j];
    const size_t chunk = (nn + gridDim.x - 1) / gridDim.x;
    const size_t lo = (size_t)blockIdx.x * chunk, hi = lo + chunk < nn ? lo + chunk : nn;
    const double* __restrict__ x = g.x[j];
    double* pp = part + (size_t)(g.job0 + j) * 8 * kDotBlocks;
    for (int r = 0; r < g.m[j]; ++r) {
        const double* __restrict__ y = g.y[j][r];
        const int kind = g.kind[j][r];
        double acc = 0;
        if (kind == 0) {
            for (size_t i = lo + threadIdx.x; i < hi; i += blockDim.x) acc += x[i] * y[i];
        } else if (kind == 1) {
            for (size_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
                const double d = x[i] - y[i];
                acc += d * d;
            }
        } else {
            for (size_t i = lo + threadIdx.x; i < hi; i += blockDim.x) acc = fmax(acc, fabs(x[i] - y[i]));
        }
        sh[threadIdx.x] = acc;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s)
                sh[threadIdx.x] = kind == 2 ? fmax(sh[threadIdx.x], sh[threadIdx.x + s]) : sh[threadIdx.x] + sh[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) pp[r * kDotBlocks + blockIdx.x] = sh[0];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(64) bred_final_kernel(BRed g, const double* __restrict__ part,
                                                        double* __restrict__ out) {
    const int j = blockIdx.x;
    const double* pp = part + (size_t)(g.job0 + j) * 8 * kDotBlocks;
    for (int r = 0; r < g.m[j]; ++r) {
        double v[1] = {pp[r * kDotBlocks + threadIdx.x]};
        if (g.kind[j][r] == 2) {
            for (int s = 32; s >= 1; s >>= 1) v[0] = fmax(v[0], __shfl_xor(v[0], s));
        } else {
            hfgpu::Halving<1, 32>::run(v, (int)threadIdx.x);
        }
        if (threadIdx.x == 0) out[(size_t)(g.job0 + j) * 8 + r] = v[0];
    }
}

// ---------------------------------------------------------------- SP2 purification
// (Niklasson, PRB 66, 155115 (2002)): for the orthogonalised Fock matrix F' the aufbau density is the
// spectral projector onto its nocc lowest eigenvectors.  From X0 = (hi I - F')/(hi - lo), [lo, hi] the
// Gershgorin interval, every step replaces X by X^2 or 2X - X^2, whichever keeps trace(X) closer to nocc;
// stopping rule of Kruchinina, Rudberg and Rubensson, JCTC 12, 5788 (2016): once e_k = tr(X - X^2) < 1e-4,
// stop as soon as e_k >= e_{k-2}.  res[job] = the final step (the projector is X_final), -1 while running.
struct BSp2 {
    int count, s;
    int idx[kJ];               // the job's index into hist, part and res
    int tile_off[kJ + 1];
    int n[kJ];
    double nocc[kJ];
    const double* F[kJ];       // orthogonalised Fock matrix
    double* X[kJ][2];          // X_s in X[s & 1]
    double* W[kJ][2];          // X_s^2 in W[s & 1]
};

__global__ void __launch_bounds__(1024) bsp2_bounds_kernel(BSp2 g, double* __restrict__ hist, double* __restrict__ res) {
    __shared__ double part[1024];
    const int j = blockIdx.x;
    double lo, hi;
    hfgpu::gershgorin_block(g.F[j], g.n[j], part, lo, hi);
    if (threadIdx.x == 0) {
        if (hi - lo < 1e-12) hi = lo + 1.0;
        double* h = hist + (size_t)g.idx[j] * kHistStride;
        h[0] = lo;
        h[1] = hi;
        h[2] = -1.0;
        res[g.idx[j]] = -1.0;
    }
}

// Operand loads of one wavefront's k strip, issued before the step's decision is read (the decision
// needs the previous step's partial traces): up to kSp2Trips x 8 k values.
constexpr int kSp2Trips = ((kSmallDense + 15) / 16 * 4 + 31) / 32;

// Step s of every job: forms X_s from X_{s-1} and X_{s-1}^2 (or from F' and its bounds for s = 0) on the
// fly, writes X_s, and multiplies X_s X_s on the matrix cores; wavefronts on diagonal tiles write partial
// traces for the next step's decision.  Step s >= 1 first decides, identically in every wavefront, for the
// previous matrix: e_{s-1}, square, stop_{s-1} = stop_{s-2} or (e_{s-1} < 1e-4 and e_{s-1} >= e_{s-3}).  If
// stop_{s-2} held, X_{s-1} was final and the step only carries the flag forward; if stop_{s-1} holds, X_s
// is written and is the result (final = s).
// hist per job: [0] lo, [1] hi, [2] final step, then per step s at 8 + 2 s: e_s, stop_s.  part per job:
// per step s at 2 nt s: nt partial traces of X_s, then nt of X_s^2.
__global__ void __launch_bounds__(256) bsp2_step_kernel(BSp2 g, double* __restrict__ hist_all,
                                                        double* __restrict__ part_all, double* __restrict__ res) {
    __shared__ v4d part4[3][64];
    const int j = job_of_block(g.tile_off, g.count);
    // a job whose purification stopped at an earlier step (its final step is set) has nothing left to do: leave
    // before the operand loads.  (Only earlier steps: the step that stops writes the final step while the other
    // blocks of its launch still have to write their tiles of the result.)
    {
        const double fin = res[g.idx[j]];
        if (fin >= 0 && fin < g.s) return;
    }
    const int n = g.n[j], nt = (n + 15) / 16, s = g.s, t = (int)blockIdx.x - g.tile_off[j];
    // block t takes tile (tx, ty), tx <= ty, of the upper triangle (column by column); the product's element
    // (i, j) = sum_k X(i, k) X(j, k) is the same sum of the same products as (j, i), so a block writes its tile and
    // the mirrored one
    int ty = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((ty + 1) * (ty + 2) / 2 <= t) ++ty;
    while (ty * (ty + 1) / 2 > t) --ty;
    const int tx = t - ty * (ty + 1) / 2;
    double* __restrict__ hist = hist_all + (size_t)g.idx[j] * kHistStride;
    double* __restrict__ gpart = part_all + (size_t)g.idx[j] * kPartStride;
    const double nocc = g.nocc[j];
    const double* __restrict__ Fp = g.F[j];
    const double* __restrict__ Xp = g.X[j][(s + 1) & 1];
    const double* __restrict__ X2p = g.W[j][(s + 1) & 1];
    double* __restrict__ Xo = g.X[j][s & 1];
    double* __restrict__ X2o = g.W[j][s & 1];
    const int l = threadIdx.x & 63, lr = l & 15, lk = l >> 4, w = threadIdx.x >> 6;
    const int i0 = tx * 16, j0 = ty * 16;
    const bool mirror = tx != ty;
    const bool lead = t == 0 && threadIdx.x == 0;
    int kb, ke;
    hfgpu::k_strip(n, w, kb, ke);
    // raw operands: element (i, kk) of A and (j, kk) of B (X_s symmetric), from X_{s-1} and
    // X_{s-1}^2, or from F' for s = 0
    double ra[kSp2Trips][8], ra2[kSp2Trips][8], rb[kSp2Trips][8], rb2[kSp2Trips][8];
    const double* P1 = s == 0 ? Fp : Xp;
    const double* P2 = s == 0 ? Fp : X2p;
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
        const double* pp = gpart + (size_t)2 * nt * (s - 1);
        double t1 = 0, t2 = 0;
        for (int i = 0; i < nt; ++i) {
            t1 += pp[i];
            t2 += pp[nt + i];
        }
        const double e = t1 - t2;
        square = fabs(t2 - nocc) <= fabs(2.0 * t1 - t2 - nocc);
        const bool stop_pp = s >= 2 && hist[8 + 2 * (s - 2) + 1] != 0;
        if (stop_pp) {                                         // X_{s-1} already final
            if (lead) hist[8 + 2 * (s - 1) + 1] = 1.0;
            return;
        }
        const double e3 = s >= 3 ? hist[8 + 2 * (s - 3)] : 1e300;
        const bool stop_p = e < 1e-4 && e >= e3;
        if (lead) {
            hist[8 + 2 * (s - 1)] = e;
            hist[8 + 2 * (s - 1) + 1] = stop_p ? 1.0 : 0.0;
            if (stop_p) {
                hist[2] = s;
                res[g.idx[j]] = s;
            }
        }
        if (stop_p) {
            // X_s is the result: write this tile of it (and the mirrored one) and stop
            if (w == 0)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = i0 + lk + 4 * r, jj = j0 + lr;
                    if (i < n && jj < n) {
                        const size_t o = i + (size_t)jj * n;
                        Xo[o] = square ? X2p[o] : 2.0 * Xp[o] - X2p[o];
                        if (mirror) {
                            const size_t om = jj + (size_t)i * n;
                            Xo[om] = square ? X2p[om] : 2.0 * Xp[om] - X2p[om];
                        }
                    }
                }
            return;
        }
    }
    const double lo = hist[0], hi = hist[1], inv = 1.0 / (hi - lo);
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
    acc = hfgpu::reduce_waves(acc, part4);
    if (w != 0) return;
    auto xval = [&](int i, int jj) -> double {                  // element (i, jj) of X_s, zero outside
        if (i >= n || jj >= n) return 0.0;
        const size_t o = i + (size_t)jj * n;
        if (s == 0) return ((i == jj ? hi : 0.0) - Fp[o]) * inv;
        return square ? X2p[o] : 2.0 * Xp[o] - X2p[o];
    };
    double d1 = 0, d2 = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = i0 + lk + 4 * r, jj = j0 + lr;
        if (i < n && jj < n) {
            const size_t o = i + (size_t)jj * n;
            const double x = xval(i, jj);
            Xo[o] = x;
            X2o[o] = acc[r];
            if (mirror) {
                const size_t om = jj + (size_t)i * n;
                Xo[om] = xval(jj, i);
                X2o[om] = acc[r];
            }
            if (i == jj) {
                d1 += x;
                d2 += acc[r];
            }
        }
    }
    if (i0 == j0) {   // diagonal tile: partial traces, reduced over the wavefront in a fixed order
        double V[2] = {d1, d2};
        hfgpu::Halving<2, 32>::run(V, l);     // lanes 0..31 now hold the d1 total, lanes 32..63 the d2 total
        double* pp = gpart + (size_t)2 * nt * s;
        if (l == 0) pp[tx] = V[0];
        if (l == 32) pp[nt + tx] = V[0];
    }
}

// SP2 for n <= kSp2LocalMax, one workgroup per job: X stays in LDS for every step, so the whole purification is
// one launch with no host round trip.  Same steps and stopping rule as bsp2_step_kernel; the 16 wavefronts share
// the 16 x 16 tiles on and above the diagonal of each product X X on the matrix cores and keep their tiles of X^2
// in registers: once the step's traces decide between X^2 and 2X - X^2, each wavefront writes the new X over its own
// tiles and their mirror images (element (i, j) of X X is the same sum of the same products as (j, i)), so X^2
// never goes to memory and only X (n^2 doubles) needs the LDS.  The projector goes to X[j][0]; res[job] =
// the final step (-1 if the rule was not met within max_steps).
constexpr int kSp2LocalWaves = 16;
constexpr int kSp2LocalTiles = ((kSp2LocalMax + 15) / 16 * ((kSp2LocalMax + 15) / 16 + 1) / 2 + kSp2LocalWaves - 1) /
                               kSp2LocalWaves;   // tiles per wavefront, at most

__global__ void __launch_bounds__(1024) bsp2_local_kernel(BSp2 g, int max_steps, double* __restrict__ res) {
    extern __shared__ double lds[];
    const int job = blockIdx.x;
    const int n = g.n[job];
    const double nocc = g.nocc[job];
    const double* __restrict__ Fp = g.F[job];
    double* __restrict__ Xout = g.X[job][0];
    double* X = lds;              // n x n, column-major (symmetric)
    __shared__ double red[2][kSp2LocalWaves];
    __shared__ double e_hist[2];
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int l = tid & 63, lr = l & 15, lk = l >> 4, w = tid >> 6;
    double lo, hi;
    hfgpu::gershgorin_block(Fp, n, lds, lo, hi);      // (X is not set up yet)
    if (hi - lo < 1e-12) hi = lo + 1.0;
    const double inv = 1.0 / (hi - lo);
    __syncthreads();
    for (int o = tid; o < n * n; o += nthr) X[o] = ((o / n == o % n ? hi : 0.0) - Fp[o]) * inv;
    if (tid == 0) e_hist[0] = e_hist[1] = 1e300;   // e_{s-1}, e_{s-2}
    __syncthreads();
    const int nt = (n + 15) / 16, ntiles = nt * (nt + 1) / 2;
    int final_step = -1;
    for (int s = 0; s <= max_steps; ++s) {
        // this wavefront's tiles of X X (tx <= ty, column by column), kept in registers
        v4d acc[kSp2LocalTiles];
        double d2 = 0.0;
#pragma unroll
        for (int q = 0; q < kSp2LocalTiles; ++q) {
            acc[q] = v4d{0.0, 0.0, 0.0, 0.0};
            const int tile = w + q * kSp2LocalWaves;
            if (tile >= ntiles) continue;
            int ty = 0;
            while ((ty + 1) * (ty + 2) / 2 <= tile) ++ty;
            const int tx = tile - ty * (ty + 1) / 2;
            const int i0 = tx * 16, j0 = ty * 16;
            v4d a4 = {0.0, 0.0, 0.0, 0.0};
            for (int k0 = 0; k0 < n; k0 += 4) {
                const int kk = k0 + lk, ia = i0 + lr, jb = j0 + lr;
                const double a = kk < n && ia < n ? X[ia + kk * n] : 0.0;
                const double b = kk < n && jb < n ? X[jb + kk * n] : 0.0;
                a4 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, a4, 0, 0, 0);
            }
            acc[q] = a4;
            if (tx == ty)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = i0 + lk + 4 * r, jj = j0 + lr;
                    if (i == jj && i < n) d2 += a4[r];
                }
        }
        double d1 = 0.0;
        for (int i = tid; i < n; i += nthr) d1 += X[i + i * n];
        {
            double V[2] = {d1, d2};
            hfgpu::Halving<2, 32>::run(V, l);  // lanes 0..31: the d1 total, lanes 32..63: the d2 total
            if (l == 0) red[0][w] = V[0];
            if (l == 32) red[1][w] = V[0];
        }
        __syncthreads();                        // every wavefront's reads of X for this step are done
        double t1 = 0, t2 = 0;
        for (int q = 0; q < kSp2LocalWaves; ++q) {
            t1 += red[0][q];
            t2 += red[1][q];
        }
        const double e = t1 - t2;
        const bool square = fabs(t2 - nocc) <= fabs(2.0 * t1 - t2 - nocc);
        const bool stop = e < 1e-4 && e >= e_hist[1];
        // X <- X^2 or 2X - X^2 over this wavefront's tiles and their mirror images (X symmetric: the old mirrored
        // element is the old element).  In a diagonal tile both (i, j) and (j, i) are the tile's own: only the lane
        // with i <= j updates the pair, so no lane reads an element another lane has already overwritten.
#pragma unroll
        for (int q = 0; q < kSp2LocalTiles; ++q) {
            const int tile = w + q * kSp2LocalWaves;
            if (tile >= ntiles) continue;
            int ty = 0;
            while ((ty + 1) * (ty + 2) / 2 <= tile) ++ty;
            const int tx = tile - ty * (ty + 1) / 2;
            const int i0 = tx * 16, j0 = ty * 16;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + lk + 4 * r, jj = j0 + lr;
                if (i < n && jj < n && (tx != ty || i <= jj)) {
                    const double x = X[i + jj * n];
                    const double xn = square ? acc[q][r] : 2.0 * x - acc[q][r];
                    X[i + jj * n] = xn;
                    X[jj + i * n] = xn;
                }
            }
        }
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
    if (tid == 0) res[g.idx[job]] = final_step;
}

__global__ void fill_kernel(double* __restrict__ p, int count, double v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

int grid_x(size_t cnt) {
    const size_t b = (cnt + 255) / 256;
    return (int)std::min<size_t>(std::max<size_t>(b, 1), 4096);
}

// ---------------------------------------------------------------- launch helpers (split into kJ jobs)
struct GemmJob {
    int n;
    double alpha;
    const double* A;
    const double* B;
    double* C;
};

hipError_t launch_gemms(hipStream_t st, const std::vector<GemmJob>& jobs) {
    for (size_t j0 = 0; j0 < jobs.size(); j0 += kJ) {
        BGemm g{};
        g.count = (int)std::min<size_t>(kJ, jobs.size() - j0);
        int tiles = 0;
        for (int j = 0; j < g.count; ++j) {
            const GemmJob& q = jobs[j0 + j];
            g.tile_off[j] = tiles;
            g.n[j] = q.n;
            g.alpha[j] = q.alpha;
            g.A[j] = q.A;
            g.B[j] = q.B;
            g.C[j] = q.C;
            const int nt = (q.n + 15) / 16;
            tiles += nt * nt;
        }
        g.tile_off[g.count] = tiles;
        hipLaunchKernelGGL(bgemm16_kernel, dim3(tiles), dim3(256), 0, st, g);
    }
    return hipGetLastError();
}

struct ElemJob {
    int n = 0;
    size_t cnt = 0;
    double x = 0;
    int flag = 0;
    const double* a = nullptr;
    const double* b = nullptr;
    const double* c = nullptr;
    double* out = nullptr;
    double* out2 = nullptr;
};

hipError_t launch_elem(hipStream_t st, int op, const std::vector<ElemJob>& jobs) {
    for (size_t j0 = 0; j0 < jobs.size(); j0 += kJ) {
        BElem g{};
        g.count = (int)std::min<size_t>(kJ, jobs.size() - j0);
        g.op = op;
        size_t mx = 1;
        for (int j = 0; j < g.count; ++j) {
            const ElemJob& q = jobs[j0 + j];
            g.n[j] = q.n;
            g.cnt[j] = q.cnt;
            g.x[j] = q.x;
            g.flag[j] = q.flag;
            g.a[j] = q.a;
            g.b[j] = q.b;
            g.c[j] = q.c;
            g.out[j] = q.out;
            g.out2[j] = q.out2;
            mx = std::max(mx, q.cnt);
        }
        hipLaunchKernelGGL(belem_kernel, dim3(grid_x(mx), g.count), dim3(256), 0, st, g);
    }
    return hipGetLastError();
}

struct RedJob {
    size_t nn;
    const double* x;
    const double* y[6];
    unsigned char kind[6];
    int m;
};

// Results of job q at out[q * 8 + r].
hipError_t launch_reductions(hipStream_t st, const std::vector<RedJob>& jobs, double* part, double* out) {
    for (size_t j0 = 0; j0 < jobs.size(); j0 += kJ) {
        BRed g{};
        g.count = (int)std::min<size_t>(kJ, jobs.size() - j0);
        g.job0 = (int)j0;
        for (int j = 0; j < g.count; ++j) {
            const RedJob& q = jobs[j0 + j];
            g.nn[j] = q.nn;
            g.x[j] = q.x;
            g.m[j] = q.m;
            for (int r = 0; r < q.m; ++r) {
                g.y[j][r] = q.y[r];
                g.kind[j][r] = q.kind[r];
            }
        }
        hipLaunchKernelGGL(bred_partial_kernel, dim3(kDotBlocks, g.count), dim3(256), 0, st, g, part);
        hipLaunchKernelGGL(bred_final_kernel, dim3(g.count), dim3(64), 0, st, g, part, out);
    }
    return hipGetLastError();
}

struct LinJob {
    size_t nn;
    const double* base;
    double* out;
    std::vector<int> slots;
    std::vector<double> c;
};

hipError_t launch_lincomb(hipStream_t st, const std::vector<LinJob>& jobs) {
    for (size_t j0 = 0; j0 < jobs.size(); j0 += kJ) {
        BLin g{};
        g.count = (int)std::min<size_t>(kJ, jobs.size() - j0);
        size_t mx = 1;
        for (int j = 0; j < g.count; ++j) {
            const LinJob& q = jobs[j0 + j];
            g.nn[j] = q.nn;
            g.base[j] = q.base;
            g.out[j] = q.out;
            g.m[j] = (int)q.slots.size();
            for (size_t r = 0; r < q.slots.size(); ++r) {
                g.slot[j][r] = (unsigned char)q.slots[r];
                g.c[j][r] = q.c[r];
            }
            mx = std::max(mx, q.nn);
        }
        hipLaunchKernelGGL(blin_kernel, dim3(grid_x(mx), g.count), dim3(256), 0, st, g);
    }
    return hipGetLastError();
}

struct Sp2Job {
    int n;
    int nocc;
    const double* F;
    double* X0;
    double* X1;
    double* W0;
    double* W1;
};

// Descriptor of jobs[pick[j0 .. j0 + kJ)] (global indices base + pick[.]); tiles: the stepped kernel's tile table.
BSp2 sp2_desc(const std::vector<Sp2Job>& jobs, const std::vector<int>& pick, size_t j0, int base, bool tiles) {
    BSp2 g{};
    g.count = (int)std::min<size_t>(kJ, pick.size() - j0);
    int t = 0;
    for (int j = 0; j < g.count; ++j) {
        const int q = pick[j0 + j];
        const Sp2Job& b = jobs[q];
        g.idx[j] = base + q;
        g.tile_off[j] = t;
        g.n[j] = b.n;
        g.nocc[j] = b.nocc;
        g.F[j] = b.F;
        g.X[j][0] = b.X0;
        g.X[j][1] = b.X1;
        g.W[j][0] = b.W0;
        g.W[j][1] = b.W1;
        if (tiles) {
            const int nt = (b.n + 15) / 16;
            t += nt * (nt + 1) / 2;     // the tiles on and above the diagonal
        }
    }
    g.tile_off[g.count] = t;
    return g;
}

// Column-pivoted Householder QR solve of the (m x m) DIIS system (row-major A), as the reference's
// Eigen colPivHouseholderQr().solve (HartreeFock/DIIS.h:60-68).
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

// DIIS extrapolation coefficients of one spin's history (B c = [0..0 1] over the kept slots, oldest first)
// and the error estimate sqrt(B_last,last + B_prev,prev) (HartreeFock/DIIS.h:40-68).
void diis_coefficients(const Diis& d, LinJob& lj, double& est) {
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
    lj.slots.assign(d.order.begin(), d.order.end());
    lj.c.assign(c.begin(), c.begin() + m);
}

int readback(hf_ctx* ctx, hipStream_t st, BatchScratch& bs, const double* src, size_t count) {
    HF_HIP_CHECK(ctx, hipMemcpyAsync(bs.hpin, src, count * sizeof(double), hipMemcpyDeviceToHost, st));
    HF_HIP_CHECK(ctx, hipStreamSynchronize(st));
    return HF_OK;
}

}  // namespace

BatchScratch::~BatchScratch() {
    if (hpin) (void)hipHostFree(hpin);
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    if (side) (void)hipStreamDestroy(side);
}

int BatchScratch::side_stream(hf_ctx* ctx) {
    if (side) return HF_OK;
    HF_HIP_CHECK(ctx, hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    for (hipEvent_t& e : ev) HF_HIP_CHECK(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return HF_OK;
}

int BatchScratch::reserve(hf_ctx* ctx, int jobs) {
    if (jobs <= cap) return HF_OK;
    const int c = std::max(jobs, 2);
    HF_HIP_CHECK(ctx, part.alloc((size_t)c * 8 * kDotBlocks));
    HF_HIP_CHECK(ctx, out.alloc((size_t)c * 8));
    HF_HIP_CHECK(ctx, hist.alloc((size_t)c * kHistStride));
    HF_HIP_CHECK(ctx, spart.alloc((size_t)c * kPartStride));
    HF_HIP_CHECK(ctx, res.alloc((size_t)c));
    if (hpin) (void)hipHostFree(hpin);
    hpin = nullptr;
    HF_HIP_CHECK(ctx, hipHostMalloc((void**)&hpin, (size_t)c * 8 * sizeof(double), hipHostMallocDefault));
    cap = c;
    return HF_OK;
}

bool sp2_local_ready() {
    // The dynamic-LDS limit of the single-workgroup kernel is an attribute of the kernel for the whole
    // process, not of one launch: set once, to the size for n = kSp2LocalMax, before any launch (setting it
    // per call from the concurrent contexts of a server changed it under the other threads' launches).
    static std::once_flag once;
    static bool ok = false;
    std::call_once(once, [] {
        const size_t lds = (size_t)kSp2LocalMax * kSp2LocalMax * sizeof(double);
        ok = hipFuncSetAttribute((const void*)bsp2_local_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)lds) == hipSuccess;
        (void)hipGetLastError();
    });
    return ok;
}

int scf_gemm(hf_ctx* ctx, hipStream_t st, bool ta, bool tb, int m, int n, int k, double alpha, const double* A,
             const double* B, double beta, double* C) {
    if (std::max(m, std::max(n, k)) <= kSmallDense)
        return hf_dgemm_small(ctx, st, ta, tb, m, n, k, alpha, A, ta ? k : m, B, tb ? n : k, beta, C, m);
    if (st != ctx->stream) rocblas_set_stream(ctx->blas, st);
    const rocblas_status s = rocblas_dgemm(ctx->blas, ta ? rocblas_operation_transpose : rocblas_operation_none,
                                           tb ? rocblas_operation_transpose : rocblas_operation_none, m, n, k, &alpha,
                                           A, ta ? k : m, B, tb ? n : k, &beta, C, m);
    if (st != ctx->stream) rocblas_set_stream(ctx->blas, ctx->stream);
    if (s != rocblas_status_success) return ctx->fail(HF_E_HIP, "rocblas_dgemm: status " + std::to_string((int)s));
    return HF_OK;
}

// ---------------------------------------------------------------- the step phases
namespace {

double* Fk(ScfRun* m, int k) { return m->ctx->scf_ws->st.Fa.p + k * m->nn; }
ScfState& SS(ScfRun* m) { return m->ctx->scf_ws->st; }

bool live(ScfRun* m) { return m->status == HF_OK; }

// Phase 1: Fock matrices.  Iteration 0 takes the guess; later iterations build J/K of the current density
// and assemble F = h + J - K/2 (RHF, symmetrised) or F_s = h + J - K_s (UHF) in one launch.  The J/K builds:
// one launch forms every incremental build's D - Dprev, the direct builds run per molecule, the in-core builds
// share one set-up launch (hf_incore_jk_batch), and one launch keeps every D as the next Dprev.  J/K time (HIP
// events) is taken only for a batch of one.
int phase_fock(const std::vector<ScfRun*>& ms, hipStream_t st) {
    std::vector<ElemJob> fock, copy, sub, keep;
    std::vector<ScfRun*> built;
    std::vector<JkPlan> plans;
    for (ScfRun* m : ms) {
        ScfState& S = SS(m);
        for (int k = 0; k < m->nsp; ++k) {
            ElemJob e;
            e.n = m->n;
            e.cnt = m->nn;
            e.out = Fk(m, k);
            if (m->step_iter == 0) {
                e.a = S.F0.p + k * m->nn;
                copy.push_back(e);
            } else {
                e.a = S.h.p;
                e.b = S.J.p;
                e.c = S.K.p + k * m->nn;
                e.x = m->uhf ? 1.0 : 0.5;
                e.flag = m->uhf ? 0 : 1;
                fock.push_back(e);
            }
        }
        if (m->step_iter == 0) continue;
        JkPlan pl;
        m->jk_plan(pl);
        ElemJob e;
        e.n = m->n;
        e.cnt = m->nsp * m->nn;
        if (pl.sub) {
            e.a = S.D.p;
            e.b = S.Dprev.p;
            e.out = S.dD.p;
            sub.push_back(e);
        }
        if (pl.copy_prev) {
            ElemJob c = e;
            c.a = S.D.p;
            c.b = nullptr;
            c.out = S.Dprev.p;
            keep.push_back(c);
        }
        built.push_back(m);
        plans.push_back(pl);
    }
    const bool timed = ms.size() == 1 && !built.empty();
    if (launch_elem(st, kOpCopy, copy) != hipSuccess) return HF_E_HIP;
    for (size_t q = 0; q < built.size(); ++q)
        if (plans[q].wait_store && hipStreamWaitEvent(st, built[q]->eri_ev[1], 0) != hipSuccess) return HF_E_HIP;
    if (timed && hipEventRecord(built[0]->jk_ev[0], st) != hipSuccess) return HF_E_HIP;
    if (launch_elem(st, kOpSub, sub) != hipSuccess) return HF_E_HIP;
    std::vector<IncoreJob> incore;
    for (size_t q = 0; q < built.size(); ++q) {
        ScfRun* m = built[q];
        const JkPlan& pl = plans[q];
        ScfState& S = SS(m);
        if (pl.kind == JkPlan::kIncore) {
            incore.push_back({m->ctx, m->nsp, m->ctx->d_G.p, pl.g32, pl.src, S.J.p, S.K.p, pl.accumulate, true});
        } else if (pl.kind == JkPlan::kDirect) {
            if (int rc = hf_jk_device_impl(m->ctx, m->nsp, pl.src, S.J.p, S.K.p, st, pl.thr, pl.accumulate, pl.fp32))
                m->status = rc;
        }
    }
    if (int rc = hf_incore_jk_batch(incore, st)) {
        if (rc == HF_E_HIP) return rc;
        for (ScfRun* m : built)          // a refused build (argument or size): those molecules' steps fail
            for (const IncoreJob& q : incore)
                if (m->ctx == q.ctx) m->status = rc;
    }
    if (launch_elem(st, kOpCopy, keep) != hipSuccess) return HF_E_HIP;
    if (timed) {
        if (hipEventRecord(built[0]->jk_ev[1], st) != hipSuccess) return HF_E_HIP;
        built[0]->jk_pending = true;
    }
    for (size_t q = 0; q < built.size(); ++q) built[q]->jk_done(plans[q]);
    // Fock matrices of the molecules whose build did not fail
    std::vector<ElemJob> fock_ok;
    size_t f = 0;
    for (ScfRun* m : ms) {
        if (m->step_iter == 0) continue;
        for (int k = 0; k < m->nsp; ++k, ++f)
            if (live(m)) fock_ok.push_back(fock[f]);
    }
    return launch_elem(st, kOpFock, fock_ok) == hipSuccess ? HF_OK : (int)HF_E_HIP;
}

// Phase 2: DIISStep (HartreeFock/RestrictedHartreeFock.cpp:91-120): e = S D F - F D S into the history with
// F, inner products of the new e with the kept ones (one read-back for the batch), the DIIS systems on the
// host, and the extrapolated Fock matrices in one launch.
int phase_diis(const std::vector<ScfRun*>& ms, hipStream_t st, BatchScratch& bs) {
    struct DJ {
        ScfRun* m;
        int k;
        double *Ta, *Tb;
        int slot;
    };
    std::vector<DJ> dj;
    for (ScfRun* m : ms) {
        if (!live(m)) continue;
        m->diis_used = false;
        m->diis_on = m->useDIIS && m->step_iter && m->step_iter < m->p->max_diis_iterations;
        if (!m->diis_on) {
            m->lastErr = 0;
            continue;
        }
        ScfState& S = SS(m);
        for (int k = 0; k < m->nsp; ++k) dj.push_back({m, k, k ? S.T3.p : S.T1.p, k ? S.err.p : S.T2.p, 0});
    }
    if (dj.empty()) return HF_OK;
    std::vector<GemmJob> g1, g2;
    for (const DJ& d : dj) {
        ScfState& S = SS(d.m);
        g1.push_back({d.m->n, 1.0, S.S.p, S.D.p + d.k * d.m->nn, d.Ta});
        g2.push_back({d.m->n, 1.0, d.Ta, Fk(d.m, d.k), d.Tb});
    }
    hipError_t e = launch_gemms(st, g1);
    if (e == hipSuccess) e = launch_gemms(st, g2);
    std::vector<ElemJob> anti;
    std::vector<RedJob> dots;
    for (DJ& d : dj) {
        Diis& h = d.m->diis[d.k];
        d.slot = h.next_slot();
        if ((int)h.order.size() == Diis::kKeep) h.order.pop_front();
        h.order.push_back(d.slot);
        ElemJob a;
        a.n = d.m->n;
        a.cnt = d.m->nn;
        a.a = d.Tb;
        a.b = Fk(d.m, d.k);
        a.out = h.E + d.slot * d.m->nn;
        a.out2 = h.F + d.slot * d.m->nn;
        anti.push_back(a);
        RedJob r{};
        r.nn = d.m->nn;
        r.x = h.E + d.slot * d.m->nn;
        r.m = (int)h.order.size();
        for (int q = 0; q < r.m; ++q) {
            r.y[q] = h.E + h.order[q] * d.m->nn;
            r.kind[q] = 0;
        }
        dots.push_back(r);
    }
    if (e == hipSuccess) e = launch_elem(st, kOpAntisymCopy, anti);
    if (e == hipSuccess) e = launch_reductions(st, dots, bs.part.p, bs.out.p);
    if (e != hipSuccess) return HF_E_HIP;
    if (int rc = readback(ms[0]->ctx, st, bs, bs.out.p, dj.size() * 8)) return rc;
    for (size_t q = 0; q < dj.size(); ++q) {
        Diis& h = dj[q].m->diis[dj[q].k];
        for (int r = 0; r < (int)h.order.size(); ++r)
            h.B[dj[q].slot][h.order[r]] = h.B[h.order[r]][dj[q].slot] = bs.hpin[q * 8 + r];
    }
    std::vector<LinJob> lin;
    for (ScfRun* m : ms) {
        if (!live(m) || !m->diis_on) continue;
        if ((int)m->diis[m->nsp - 1].order.size() < Diis::kFirst) continue;
        double e0 = 0, e1 = 0;
        for (int k = m->nsp - 1; k >= 0; --k) {
            LinJob lj;
            lj.nn = m->nn;
            lj.base = m->diis[k].F;
            lj.out = Fk(m, k);
            diis_coefficients(m->diis[k], lj, k == m->nsp - 1 ? e1 : e0);
            lin.push_back(std::move(lj));
        }
        m->lastErr = e1 + e0;
        m->diis_used = true;
    }
    return launch_lincomb(st, lin) == hipSuccess ? HF_OK : (int)HF_E_HIP;
}

// Phase 3, for spin k: F' = X F X, its aufbau projector P (SP2: one workgroup per molecule for small n,
// one launch per purification step over all molecules otherwise), and D_new = scale X P X; a molecule
// whose purification does not resolve (no HOMO-LUMO gap) diagonalises F' instead, as the reference does
// every step (HartreeFock/RestrictedHartreeFock.cpp:122-140).
int phase_density(const std::vector<ScfRun*>& ms, int k, hipStream_t st, BatchScratch& bs) {
    std::vector<ScfRun*> mk;
    for (ScfRun* m : ms)
        if (live(m) && m->nsp > k) mk.push_back(m);
    if (mk.empty()) return HF_OK;
    hf_ctx* ctx0 = mk[0]->ctx;
    std::vector<GemmJob> g1, g2;
    std::vector<ScfRun*> big;        // n > kSmallDense: library products, one molecule at a time
    for (ScfRun* m : mk) {
        ScfState& S = SS(m);
        if (m->n > kSmallDense) {
            big.push_back(m);
            continue;
        }
        g1.push_back({m->n, 1.0, Fk(m, k), S.X.p, S.T1.p});
        g2.push_back({m->n, 1.0, S.X.p, S.T1.p, S.T2.p});
    }
    if (launch_gemms(st, g1) != hipSuccess || launch_gemms(st, g2) != hipSuccess) return HF_E_HIP;
    for (ScfRun* m : big) {
        ScfState& S = SS(m);
        if (int rc = scf_gemm(m->ctx, st, false, false, m->n, m->n, m->n, 1.0, Fk(m, k), S.X.p, 0.0, S.T1.p))
            m->status = rc;
        else if (int rc2 = scf_gemm(m->ctx, st, false, false, m->n, m->n, m->n, 1.0, S.X.p, S.T1.p, 0.0, S.T2.p))
            m->status = rc2;
    }
    // classify: empty spin, local SP2, stepped SP2, library SP2, diagonalisation
    std::vector<Sp2Job> local, stepped;
    std::vector<ScfRun*> mlocal, mstep, mdiag;
    std::vector<ElemJob> zero;
    int first_check = 0;
    bool all_hint = true;
    const bool local_ok = sp2_local_ready();
    for (ScfRun* m : mk) {
        if (!live(m)) continue;
        ScfState& S = SS(m);
        m->sp2_ok[k] = false;
        m->sp2_P[k] = nullptr;
        if (m->occ[k] == 0) {
            ElemJob z;
            z.n = m->n;
            z.cnt = m->nn;
            z.out = S.Dn.p + k * m->nn;
            zero.push_back(z);
            m->sp2_ok[k] = true;
            continue;
        }
        if (m->p->density_method != 0) {
            mdiag.push_back(m);
        } else if (m->n <= kSp2LocalMax && local_ok) {
            local.push_back({m->n, m->occ[k], S.T2.p, S.T1.p, S.T3.p, nullptr, nullptr});
            mlocal.push_back(m);
        } else if (m->n <= kSmallDense) {
            stepped.push_back({m->n, m->occ[k], S.T2.p, S.T1.p, S.T3.p, S.sp2W.p, S.sp2W.p + m->nn});
            mstep.push_back(m);
            const int hint = m->sp2_steps[k];
            if (hint <= 0) all_hint = false;
            first_check = std::max(first_check, hint > 0 ? std::max(8, hint) : 16);
        } else {
            bool ok = false;
            if (int rc = scf_sp2_projector(m->ctx, S, st, S.T2.p, m->occ[k], S.T1.p, S.T3.p, ok)) {
                m->status = rc;
                continue;
            }
            m->sp2_ok[k] = ok;
            m->sp2_P[k] = ok ? S.T1.p : nullptr;
            if (!ok) mdiag.push_back(m);
        }
    }
    if (launch_elem(st, kOpZero, zero) != hipSuccess) return HF_E_HIP;
    const int nl = (int)local.size(), ns = (int)stepped.size();
    if (int rc = bs.reserve(ctx0, nl + ns)) return rc;
    // The single-workgroup purifications run beside the stepped ones, on the scratch's side stream, when there
    // are both.
    const bool fork = nl > 0 && ns > 0;
    hipStream_t ls = st;
    if (fork) {
        if (int rc = bs.side_stream(ctx0)) return rc;
        ls = bs.side;
        if (hipEventRecord(bs.ev[0], st) != hipSuccess || hipStreamWaitEvent(ls, bs.ev[0], 0) != hipSuccess)
            return HF_E_HIP;
    }
    if (nl) {
        // results start at "not met", so nothing a launch failed to write is read as a result
        hipLaunchKernelGGL(fill_kernel, dim3((nl + 255) / 256), dim3(256), 0, ls, bs.res.p, nl, -1.0);
        std::vector<int> all(nl);
        for (int q = 0; q < nl; ++q) all[q] = q;
        for (size_t j0 = 0; j0 < all.size(); j0 += kJ) {
            BSp2 g = sp2_desc(local, all, j0, 0, false);
            size_t mx = 0;
            for (int j = 0; j < g.count; ++j) mx = std::max(mx, (size_t)g.n[j] * g.n[j]);
            const size_t lds = std::max<size_t>(mx, 1024) * sizeof(double);   // >= the bounds' workspace
            hipLaunchKernelGGL(bsp2_local_kernel, dim3(g.count), dim3(1024), lds, ls, g, kSp2MaxSteps, bs.res.p);
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return ctx0->fail(HF_E_HIP, std::string("single-workgroup SP2 launch: ") + hipGetErrorString(e));
        if (fork && hipEventRecord(bs.ev[1], ls) != hipSuccess) return HF_E_HIP;
    }
    if (ns) {
        // The stop flags are read back from step first_check on (the largest step count of the molecules'
        // previous purifications, when every one has one) and then every 2 (4 without hints) steps; molecules
        // found finished are left out of the later launches (those that finish in between leave each launch
        // at once).
        const int every = all_hint ? 2 : 4;
        std::vector<int> open(ns);
        for (int q = 0; q < ns; ++q) open[q] = q;
        for (size_t j0 = 0; j0 < open.size(); j0 += kJ) {
            BSp2 g = sp2_desc(stepped, open, j0, nl, true);
            hipLaunchKernelGGL(bsp2_bounds_kernel, dim3(g.count), dim3(1024), 0, st, g, bs.hist.p, bs.res.p);
        }
        std::vector<BSp2> descs;
        auto build = [&]() {
            descs.clear();
            for (size_t j0 = 0; j0 < open.size(); j0 += kJ) descs.push_back(sp2_desc(stepped, open, j0, nl, true));
        };
        build();
        for (int s = 0; s <= kSp2MaxSteps; ++s) {
            for (BSp2& g : descs) {
                g.s = s;
                hipLaunchKernelGGL(bsp2_step_kernel, dim3(g.tile_off[g.count]), dim3(256), 0, st, g, bs.hist.p,
                                   bs.spart.p, bs.res.p);
            }
            if ((s >= first_check && (s - first_check) % every == 0) || s == kSp2MaxSteps) {
                if (hipGetLastError() != hipSuccess) return HF_E_HIP;
                if (int rc = readback(ctx0, st, bs, bs.res.p + nl, ns)) return rc;
                std::vector<int> still;
                for (int q : open)
                    if (bs.hpin[q] < 0) still.push_back(q);
                if (still.empty()) break;
                if (still.size() != open.size()) {
                    open.swap(still);
                    build();
                }
            }
        }
    }
    if (fork && hipStreamWaitEvent(st, bs.ev[1], 0) != hipSuccess) return HF_E_HIP;
    if (nl + ns) {
        if (hipGetLastError() != hipSuccess) return HF_E_HIP;
        if (int rc = readback(ctx0, st, bs, bs.res.p, nl + ns)) return rc;
        for (int q = 0; q < nl + ns; ++q) {
            ScfRun* m = q < nl ? mlocal[q] : mstep[q - nl];
            ScfState& S = SS(m);
            const double fin = bs.hpin[q];
            const bool ok = fin >= 0;
            const int steps = (int)fin;
            m->sp2_ok[k] = ok;
            m->sp2_P[k] = !ok ? nullptr : (q < nl || (steps & 1) == 0) ? S.T1.p : S.T3.p;
            m->sp2_steps[k] = ok ? steps : 0;
            m->sp2_total += ok ? steps + 1 : 0;
            if (!ok) mdiag.push_back(m);
        }
    }
    // D_new = scale X P X for the resolved molecules
    std::vector<GemmJob> g3, g4;
    for (ScfRun* m : mk) {
        if (!live(m) || m->occ[k] == 0 || !m->sp2_ok[k]) continue;
        ScfState& S = SS(m);
        const double scale = m->uhf ? 1.0 : 2.0;
        double* W = m->sp2_P[k] == S.T1.p ? S.T3.p : S.T1.p;
        double* Dk = S.Dn.p + k * m->nn;
        if (m->n > kSmallDense) {
            if (int rc = scf_gemm(m->ctx, st, false, false, m->n, m->n, m->n, 1.0, m->sp2_P[k], S.X.p, 0.0, W))
                m->status = rc;
            else if (int rc2 = scf_gemm(m->ctx, st, false, false, m->n, m->n, m->n, scale, S.X.p, W, 0.0, Dk))
                m->status = rc2;
            continue;
        }
        g3.push_back({m->n, 1.0, m->sp2_P[k], S.X.p, W});
        g4.push_back({m->n, scale, S.X.p, W, Dk});
    }
    if (launch_gemms(st, g3) != hipSuccess || launch_gemms(st, g4) != hipSuccess) return HF_E_HIP;
    // diagonalisation: F' = U e U^T, C = X U, D = scale C_occ C_occ^T
    for (ScfRun* m : mdiag) {
        if (!live(m)) continue;
        ScfState& S = SS(m);
        const double scale = m->uhf ? 1.0 : 2.0;
        ++m->diag_steps;
        int rc = scf_syevd(m->ctx, S, st, S.T2.p, S.eps.p + k * m->n);
        if (!rc) rc = scf_gemm(m->ctx, st, false, false, m->n, m->n, m->n, 1.0, S.X.p, S.T2.p, 0.0, S.C.p + k * m->nn);
        if (!rc)
            rc = scf_gemm(m->ctx, st, false, true, m->n, m->n, m->occ[k], scale, S.C.p + k * m->nn, S.C.p + k * m->nn,
                          0.0, S.Dn.p + k * m->nn);
        if (rc) m->status = rc;
    }
    // The reference's density 2 C_occ C_occ^T is symmetric to the last bit (element (i,j) and (j,i) are the
    // same products summed in the same order); X P X from tiled matrix-core products is not, and with an
    // ill-conditioned S (diffuse functions: large elements in X) the two triangles differ by ~1e-10.  The
    // J/K builds read one triangle or the other, so the density is made exactly symmetric.
    std::vector<ElemJob> sym;
    for (ScfRun* m : mk) {
        if (!live(m) || m->occ[k] == 0) continue;
        m->last_diag[k] = !m->sp2_ok[k];
        ElemJob e;
        e.n = m->n;
        e.cnt = m->nn;
        e.out = SS(m).Dn.p + k * m->nn;
        sym.push_back(e);
    }
    for (ScfRun* m : mk)
        if (live(m) && m->occ[k] == 0) m->last_diag[k] = false;
    return launch_elem(st, kOpSym, sym) == hipSuccess ? HF_OK : (int)HF_E_HIP;
}

// Phase 4: energy E = (D.h + D.F)/2 summed over spins (D.F is twice (RHF) or once (UHF) the sum of the
// occupied orbital energies; HartreeFock/RestrictedHartreeFock.cpp:234-249) and ||D_new - D||_F, one
// read-back for the batch; then D <- D_new (DIIS) or alpha D_new + (1 - alpha) D.
int phase_energy(const std::vector<ScfRun*>& ms, hipStream_t st, BatchScratch& bs) {
    std::vector<RedJob> red;
    std::vector<ScfRun*> mr;
    for (ScfRun* m : ms) {
        if (!live(m)) continue;
        ScfState& S = SS(m);
        for (int k = 0; k < m->nsp; ++k) {
            RedJob r{};
            r.nn = m->nn;
            r.x = S.Dn.p + k * m->nn;
            r.m = 4;
            r.y[0] = S.h.p;
            r.kind[0] = 0;
            r.y[1] = Fk(m, k);
            r.kind[1] = 0;
            r.y[2] = S.D.p + k * m->nn;
            r.kind[2] = 1;
            r.y[3] = S.D.p + k * m->nn;
            r.kind[3] = 2;
            red.push_back(r);
            mr.push_back(m);
        }
    }
    if (red.empty()) return HF_OK;
    hf_ctx* ctx0 = mr[0]->ctx;
    if (int rc = bs.reserve(ctx0, (int)red.size())) return rc;
    if (launch_reductions(st, red, bs.part.p, bs.out.p) != hipSuccess) return HF_E_HIP;
    if (int rc = readback(ctx0, st, bs, bs.out.p, red.size() * 8)) return rc;
    std::vector<ElemJob> mix;
    size_t q = 0;
    for (ScfRun* m : ms) {
        if (!live(m)) continue;
        double e = 0, rms2 = 0, mx = 0;
        for (int k = 0; k < m->nsp; ++k, ++q) {
            const double* r = bs.hpin + q * 8;
            e += 0.5 * (r[0] + r[1]);
            rms2 += r[2];
            mx = std::max(mx, r[3]);
        }
        m->energy = e + m->Enuc;
        m->rms = std::sqrt(rms2);
        m->dmax_next = (m->diis_used ? 1.0 : m->p->alpha) * mx;
        if (m->jk_pending) {
            float t = 0;
            if (hipEventElapsedTime(&t, m->jk_ev[0], m->jk_ev[1]) == hipSuccess) m->jk_ms += t;
            m->jk_pending = false;
        }
        ScfState& S = SS(m);
        ElemJob x;
        x.n = m->n;
        x.cnt = m->nsp * m->nn;
        x.a = S.Dn.p;
        x.out = S.D.p;
        x.x = m->p->alpha;
        x.flag = m->diis_used ? 1 : 0;
        mix.push_back(x);
    }
    return launch_elem(st, kOpMix, mix) == hipSuccess ? HF_OK : (int)HF_E_HIP;
}

}  // namespace

int scf_step_phases(const std::vector<ScfRun*>& ms, hipStream_t st, BatchScratch& bs) {
    if (ms.empty()) return HF_OK;
    for (ScfRun* m : ms) m->status = HF_OK;
    int jobs = 0;
    for (ScfRun* m : ms) jobs += m->nsp;
    if (int rc = bs.reserve(ms[0]->ctx, jobs)) return rc;
    if (int rc = phase_fock(ms, st)) return rc;
    if (int rc = phase_diis(ms, st, bs)) return rc;
    for (int k = 0; k < 2; ++k)
        if (int rc = phase_density(ms, k, st, bs)) return rc;
    return phase_energy(ms, st, bs);
}

// ---------------------------------------------------------------- lockstep SCF of several molecules
extern "C" int hf_scf_batch(hf_ctx* const* ctxs, int count, const hf_scf_params* params, hf_scf_result* results,
                            int* status) {
    if (!ctxs || count <= 0 || !params || !results || !status) return HF_E_ARG;
    for (int i = 0; i < count; ++i) {
        if (!ctxs[i]) return HF_E_ARG;
        for (int j = 0; j < i; ++j)
            if (ctxs[j] == ctxs[i]) return ctxs[i]->fail(HF_E_ARG, "a context appears twice in the batch");
    }
    hf_ctx* lead = ctxs[0];
    const hipStream_t st = lead->stream;
    std::vector<std::unique_ptr<ScfRun>> runs(count);
    std::vector<ScfRun*> active;
    std::vector<int> index;
    for (int i = 0; i < count; ++i) {
        std::memset(&results[i], 0, sizeof(results[i]));
        status[i] = HF_OK;
        hf_ctx* ctx = ctxs[i];
        if (!ctx->have_mol) {
            status[i] = ctx->fail(HF_E_STATE, "no molecule set");
            continue;
        }
        hf_scf_session_free(ctx);
        runs[i].reset(new (std::nothrow) ScfRun(ctx, params[i]));
        if (!runs[i]) {
            status[i] = ctx->fail(HF_E_OOM, "SCF session");
            continue;
        }
        // Init on the molecule's own stream; the batch's steps (on the lead context's stream) wait for it
        if (int rc = runs[i]->init()) {
            status[i] = rc;
            continue;
        }
        if (ctx->stream != st && hipStreamWaitEvent(st, ctx->scf_ws->ready_ev, 0) != hipSuccess) {
            status[i] = ctx->fail(HF_E_HIP, "hipStreamWaitEvent");
            continue;
        }
        runs[i]->begin();
        active.push_back(runs[i].get());
        index.push_back(i);
    }
    BatchScratch bs;
    auto finish_one = [&](size_t q) {
        ScfRun* m = active[q];
        const int i = index[q];
        hf_ctx* ctx = m->ctx;
        // the molecule's last step ran on the batch's stream; finish() works on its own
        if (ctx->stream != st) {
            if (hipEventRecord(ctx->scf_ws->ready_ev, st) != hipSuccess ||
                hipStreamWaitEvent(ctx->stream, ctx->scf_ws->ready_ev, 0) != hipSuccess) {
                status[i] = ctx->fail(HF_E_HIP, "stream ordering at the end of the batch");
                return;
            }
        }
        status[i] = m->finish(&results[i], nullptr);
    };
    while (!active.empty()) {
        for (ScfRun* m : active) m->step_iter = m->next_iter();
        if (int rc = scf_step_phases(active, st, bs)) {
            for (size_t q = 0; q < active.size(); ++q) {
                status[index[q]] = rc;
                if (active[q]->ctx != lead) active[q]->ctx->err = lead->err;
            }
            break;
        }
        std::vector<ScfRun*> keep;
        std::vector<int> keep_index;
        for (size_t q = 0; q < active.size(); ++q) {
            ScfRun* m = active[q];
            int rc = m->status;
            if (rc == HF_OK) rc = m->after_step();
            if (rc != HF_OK) {
                status[index[q]] = rc;
                continue;
            }
            if (m->done()) {
                finish_one(q);
                continue;
            }
            keep.push_back(m);
            keep_index.push_back(index[q]);
        }
        active.swap(keep);
        index.swap(keep_index);
    }
    for (int i = 0; i < count; ++i)
        if (status[i] != HF_OK) return status[i];
    return HF_OK;
}
