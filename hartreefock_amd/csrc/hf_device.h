// Device-side data layout shared by the host launcher code and the ERI kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace hfgpu {

// Boys table: F_m(T_k), T_k = k * BOYS_DT, k = 0..BOYS_NROW-1, m = 0..BOYS_NCOL-1.
// The kernels use an 8-term Taylor expansion about the nearest grid point for
// F_L(T) and downward recursion for lower m while T < BOYS_TMAX, and the
// asymptotic F_0 with upward recursion beyond.
constexpr int BOYS_NCOL = 16;
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

// One class of shell pairs (La, Lb), La >= Lb, sorted by Schwarz bound, descending.
struct PairClassDev {
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

enum EriMode { MODE_JK = 0, MODE_PACKED = 1, MODE_DIAG = 2 };

// The quartet tasks of one launch: bra pair p (in class `bra`) x ket pair q
// (in class `ket`).  Tasks are enumerated by a prefix array over bra pairs:
// bra p owns tasks [prefix[p], prefix[p+1]) and its ket pairs are q = 0..count-1
// of the (descending-Q) ket list.  In MODE_DIAG task t is the quartet (t|t).
struct TaskList {
    PairClassDev bra, ket;
    const int64_t* prefix;   // nbra + 1 entries (unused in MODE_DIAG)
    int64_t ntasks;
    int same_class;          // bra and ket are one class: apply the p==q degeneracy
};

struct SinkArgs {
    int n;                   // basis functions
    int ndens;               // densities contracted in one pass (1 = RHF, 2 = UHF)
    const double* D;         // ndens x n x n, row-major, symmetric
    double* J;               // n x n accumulator (J' ; final J = J' + J'^T), from sum of densities
    double* K;               // ndens x n x n accumulators (K' ; final K = K' + K'^T)
    double* packed;          // MODE_PACKED output (reference 8-fold packed order)
    uint64_t packed_len;
    double* Qout;            // MODE_DIAG output, one per bra pair
    int lds_slots;           // MODE_JK: bra pairs per block whose exchange rows are staged in LDS
};

}  // namespace hfgpu
