// Device-side data layout shared by the host launcher code and the ERI kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace hfgpu {

// Boys table: row k (T_k = k * BOYS_DT, k = 0..BOYS_NROW-1) holds F_m(T_k) for m = 0..BOYS_NCOL-1
// and e^{-T_k} in column BOYS_NCOL; rows are BOYS_STRIDE doubles apart.
// The kernels use an 8-term Taylor expansion about the nearest grid point for
// F_L(T), e^{-T} = e^{-T_k} e^{T_k - T} (Taylor for the small factor) and downward
// recursion for lower m while T < BOYS_TMAX, and the asymptotic F_0 with upward
// recursion beyond.
constexpr int BOYS_NCOL = 16;
constexpr int BOYS_STRIDE = BOYS_NCOL + 1;
constexpr int BOYS_TAYLOR = 8;
constexpr double BOYS_DT = 0.1;
constexpr double BOYS_INV_DT = 10.0;
constexpr double BOYS_TMAX = 36.0;
constexpr int BOYS_NROW = 361;

// One primitive pair of a shell pair (64 bytes).
//   P = (a A + b B)/zeta, Kz = c_a c_b N_a N_b exp(-a b/zeta |AB|^2) / zeta,
// with N the radial normalisation (2a/pi)^{3/4} (4a)^{L/2}.
struct alignas(64) PrimPair {
    double Px, Py, Pz, zeta;
    double Kz, inv_zeta, pad0, pad1;
};

// Everything a quartet kernel needs about one shell pair, in one 32-byte record.
struct alignas(32) PairRec {
    int sa, sb;           // shells (first has L = La)
    int prim0, nprim;     // primitive pairs
    int a0, b0;           // first basis function of each shell
    int pad0, pad1;
};

// One class of shell pairs (La, Lb), La >= Lb, sorted by Schwarz bound, descending.
struct PairClassDev {
    const PairRec* rec;
    const int* sa;        // first shell (L = La)
    const int* sb;        // second shell (L = Lb)
    const int* prim0;     // offset of the pair's primitive pairs
    const int* nprim;     // number of primitive pairs
    const double* Q;      // Schwarz bound sqrt(max |(ab|ab)|)
    int npairs;
};

struct ShellsDev {
    const double* xyz;    // 3 * nshell
    const int* bf;        // first basis function
};

// MODE_JK32: J/K digestion (in FP64) of quartets evaluated in FP32 (mixed-precision tier).
enum EriMode { MODE_JK = 0, MODE_PACKED = 1, MODE_DIAG = 2, MODE_JK32 = 3 };

// Within a pair class, pairs are grouped by contraction length (number of primitive pairs,
// descending) and sorted by Schwarz bound (descending) inside each group.  A segment is one bra
// group of a quartet class, so every bra in it has the same primitive-pair count.  Its tasks are
// enumerated by a prefix array over the bra group: bra p = bra0 + i owns tasks
// [prefix[i], prefix[i+1]), and those run over the ket groups in order, taking in each ket
// group g the koff[i][g+1] - koff[i][g] kets that start kbeg[i][g] places into the group (Q
// descending: the screened head, or for a band of bounds a range further in).  Consecutive
// tasks of one bra therefore share the bra and, except at a group change, the ket contraction
// length: wavefronts see little divergence in the primitive loops and the bra-side J/K
// contributions of a block stay on one or two bra pairs.
struct TaskSeg {
    int bra0, nbra, nkg, pad;
    int64_t prefix_off;      // first of the segment's nbra + 1 prefix entries
    int64_t koff_off;        // first of the segment's nbra * (nkg + 1) ket-group offsets
    int64_t kbeg_off;        // first of the segment's nbra * nkg ket-range starts
    int64_t ntasks;
    int64_t task0;           // first task of the segment within the launch
    int64_t block0;          // first 256-task block (segments padded to whole blocks)
};

// The quartet tasks of one launch: the segments of one quartet class.  In MODE_DIAG task t is
// the quartet (t|t) of the bra class and there are no segments.  `tasks` is the same list
// expanded once per task set to (bra, ket) pairs in launch order, nblocks * 256 entries (segment
// tails padded with (-1, -1)): a kernel reads its task with one load.
struct TaskList {
    PairClassDev bra, ket;
    const int2* tasks;
    const int64_t* prefix;   // prefix arrays of every segment
    const int* koff;         // per-bra ket-group offsets of every segment
    const int* kbeg;         // per-bra ket-range starts within each ket group
    const int* kgroup;       // first pair of each ket group (ket class), nkg + 1 entries
    const TaskSeg* seg;
    int nseg;
    int64_t ntasks;          // sum over segments (MODE_DIAG: bra pairs)
    int64_t nblocks;         // sum over segments of whole 256-task blocks
    int same_class;          // bra and ket are one class: apply the p==q degeneracy
    int slice_prims;         // primitive-slice target of this class's tasks (0: whole quartets)
};

// Segment containing launch unit x: the largest s with key(s) <= x, key = block0 (by_block) or task0.
__device__ __forceinline__ int segment_of(const TaskSeg* __restrict__ seg, int nseg, int64_t x, bool by_block) {
    int lo = 0, hi = nseg;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if ((by_block ? seg[mid].block0 : seg[mid].task0) <= x) lo = mid; else hi = mid;
    }
    return lo;
}

// Ket pair for rank r (0 <= r < ko[nkg]) in the ket list of one bra, whose group offsets are ko
// and range starts kb.
__device__ __forceinline__ int ket_of_rank(const int* __restrict__ ko, const int* __restrict__ kb,
                                           const int* __restrict__ kgroup, int nkg, int r) {
    int g = 0;
    while (g + 1 < nkg && ko[g + 1] <= r) ++g;
    return kgroup[g] + kb[g] + (r - ko[g]);
}

// Ket pair of segment-local task t owned by bra index pl of segment sg.
__device__ __forceinline__ int ket_of_task(const TaskList& tl, const TaskSeg& sg, const int64_t* __restrict__ pre, int pl,
                                           int64_t t) {
    return ket_of_rank(tl.koff + sg.koff_off + (int64_t)pl * (sg.nkg + 1), tl.kbeg + sg.kbeg_off + (int64_t)pl * sg.nkg,
                       tl.kgroup, sg.nkg, (int)(t - pre[pl]));
}

// Bra pair (index within the segment's bra group) owning task t: the largest i in [lo, hi) with
// prefix[i] <= t (prefix[hi] > t).
__device__ __forceinline__ int bra_of_task(const int64_t* __restrict__ prefix, int64_t t, int lo, int hi) {
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

// Primitive slices of a J/K task.  A quartet with many primitive quartets (npb x npk) is one long
// dependent chain in one thread; in a class with few quartets (a small molecule) that chain is the
// launch's duration.  J/K tasks of such a class split the bra primitives of a quartet with at
// least 1.5 x target primitive quartets into slices of about target primitive quartets, one task
// each: every slice evaluates its partial contracted block and digests it (J and K are linear in
// the integrals).  A task's ket word carries the slice + 1 above kTaskKetBits (0: whole quartet).
constexpr int kTaskKetBits = 24;
constexpr int kTaskKetMask = (1 << kTaskKetBits) - 1;
__host__ __device__ __forceinline__ int slice_chunk(int npb, int npk, int target) {   // bra primitives per slice
    const int tot = npb * npk;
    if (target <= 0 || npb < 2 || 2 * tot < 3 * target) return npb;
    int s = (tot + target - 1) / target;
    if (s > npb) s = npb;
    return (npb + s - 1) / s;
}
__host__ __device__ __forceinline__ int slice_count(int npb, int npk, int target) {
    const int c = slice_chunk(npb, npk, target);
    return (npb + c - 1) / c;
}

// In-core store index of the canonical function pair (k, l), k >= l: rows of the pair triangle padded
// to an even length (row k starts at pair_pad_base(k)), so every row and every even column offset is
// 16-byte aligned for FP64; n functions span pair_pad_base(n) positions.
__host__ __device__ __forceinline__ int64_t pair_pad_base(int64_t k) {
    return (k & 1) ? (k + 1) * (k + 1) / 2 : k * (k + 2) / 2;
}
__host__ __device__ __forceinline__ int64_t pair_pad_index(int64_t a, int64_t b) {
    return a >= b ? pair_pad_base(a) + b : pair_pad_base(b) + a;
}

struct SinkArgs {
    int n;                   // basis functions
    int ndens;               // densities contracted in one pass (1 = RHF, 2 = UHF)
    const double* D;         // ndens x n x n, row-major, symmetric
    double* J;               // n x n accumulator (J' ; final J = J' + J'^T), from sum of densities
    double* K;               // ndens x n x n accumulators (K' ; final K = K' + K'^T)
    double* packed;          // MODE_PACKED output (reference 8-fold packed order)
    uint64_t packed_len;
    uint64_t pair_stride;    // MODE_PACKED: 0 = reference packed order; > 0 = in-core store, the symmetric
                             // matrix over padded pair indices (pair_pad_index): value at
                             // [pab * pair_stride + pcd] and [pcd * pair_stride + pab]
    double* Qout;            // MODE_DIAG output, one per bra pair
    int lds_slots;           // MODE_JK: bra pairs per block whose exchange rows are staged in LDS
    const float* boys32;     // MODE_JK32: FP32 copy of the Boys table
    int tasks_per_block;     // MODE_JK, MODE_JK32: consecutive 256-task blocks per workgroup (set at launch)
    long long* fx;           // deterministic mode: fixed-point accumulators (see fx_add) of J' (n x n) then
                             // K' (ndens x n x n), two per element; null: FP64 atomics into J and K
};

#if defined(__HIP__)   // (the header is also read by host-only sources)
// Deterministic accumulation.  A contribution v is added as the 80-fractional-bit fixed-point number
// round(v 2^80), split over two 64-bit integers: fx[2 idx] += floor(v 2^40) and fx[2 idx + 1] += the
// remaining fraction times 2^40 (in [0, 2^40]).  Integer addition is associative, so the sums do not
// depend on the order in which threads add, and a build gives the same bits every time.  |v| < 2^23 per
// contribution; a sum of up to 2^23 contributions per element keeps the low word in range.  fx_value
// folds the two words back into a double.
__device__ __forceinline__ void fx_add(long long* fx, size_t idx, double v) {
    const double t = v * 0x1p40;
    const double h = floor(t);
    atomicAdd((unsigned long long*)fx + 2 * idx, (unsigned long long)(long long)h);
    atomicAdd((unsigned long long*)fx + 2 * idx + 1, (unsigned long long)(long long)rint((t - h) * 0x1p40));
}
__device__ __forceinline__ double fx_value(long long hi, long long lo) {
    const long long carry = lo >> 40;                  // lo = carry 2^40 + r, 0 <= r < 2^40
    return (double)(hi + carry) * 0x1p-40 + (double)(lo - carry * (1LL << 40)) * 0x1p-80;
}
__device__ __forceinline__ void sink_add_J(const SinkArgs& s, size_t idx, double v) {
    if (s.fx)
        fx_add(s.fx, idx, v);
    else
        unsafeAtomicAdd(&s.J[idx], v);
}
__device__ __forceinline__ void sink_add_K(const SinkArgs& s, int k, size_t idx, double v) {
    const size_t nn = (size_t)s.n * s.n;
    if (s.fx)
        fx_add(s.fx, (size_t)(k + 1) * nn + idx, v);
    else
        unsafeAtomicAdd(&s.K[(size_t)k * nn + idx], v);
}
#endif

}  // namespace hfgpu
