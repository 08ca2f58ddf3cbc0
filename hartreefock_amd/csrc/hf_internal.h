// Host-side internals of libhfgpu: molecule model, context state, helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/hfgpu.h"
#include "hf_device.h"

namespace hfhost {

// One shell after splitting SP shells: L, centre, primitives with coef = c * radial norm.
struct Shell {
    int L = 0, atom = 0, bf = 0;
    double xyz[3] = {0, 0, 0};
    std::vector<double> alpha, coef;
};

struct Molecule {
    std::vector<int> Z;
    std::vector<double> xyz;      // 3 * natoms
    std::vector<Shell> shells;
    int nbf = 0;
    int maxL = 0;
    // Molecule::alphaElectrons / betaElectrons (HartreeFock/Molecule.h:16-17): both 0 means the neutral
    // count, split by Molecule::Init (HartreeFock/Molecule.cpp:127-140)
    int alpha_set = 0, beta_set = 0;
    double field[3] = {0, 0, 0};  // Molecule::ElectricField, atomic units
    int electrons() const {
        if (alpha_set > 0 || beta_set > 0) return alpha_set + beta_set;
        int e = 0;
        for (int z : Z) e += z;
        return e;
    }
    bool has_field() const { return field[0] != 0 || field[1] != 0 || field[2] != 0; }
};

// Parse an NWChem basis file and build the molecule's shell list (basis.cpp).
bool load_molecule(const char* basis_file, int natoms, const int* Z, const double* xyz, Molecule& mol,
                   std::string& err);
bool molecule_from_soa(const hf_basis_soa* b, Molecule& mol, std::string& err);
double radial_norm(double alpha, int L);
double nuclear_repulsion(const Molecule& mol);

// Boys functions F_0..F_mmax(T) on the host (boys.cpp).
void boys_host(int mmax, double T, double* F);
// Table F_m(k*dt) for the device kernels.
void boys_table(std::vector<double>& tab);

inline int ncart(int L) { return (L + 1) * (L + 2) / 2; }

}  // namespace hfhost

// Device buffer helper (owned).
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // Grows to at least count elements.  A buffer that already holds memory grows by half again at least: every
    // hipFree synchronises the whole device (stalling the other contexts of a server), so the growth steps of a
    // context that meets larger and larger molecules are kept few.
    hipError_t alloc(size_t count) {
        if (count <= n && p) return hipSuccess;
        const size_t want = p ? std::max(count, n + n / 2) : count;
        return alloc_exact(want);
    }
    // Grows to exactly count elements (the in-core store, whose size the admission of a server charges).
    hipError_t alloc_exact(size_t count) {
        if (count <= n && p) return hipSuccess;
        release();
        if (count == 0) return hipSuccess;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e == hipSuccess) n = count; else p = nullptr;
        return e;
    }
};

struct QuartetClassTasks {
    int cls = -1;             // index into kClasses
    int bra = -1, ket = -1;   // pair-class indices
    int64_t ntasks = 0;       // tasks (a sliced quartet counts once per slice)
    int64_t quartets = 0;     // shell quartets
    int slice_prims = 0;      // primitive-slice target of the tasks (0: whole quartets)
    int64_t nblocks = 0;      // whole 256-task blocks over the segments
    int64_t prim_quartets = 0;
    size_t seg_off = 0;       // first segment in the shared device segment buffer
    int nseg = 0;
    size_t tasks_off = 0;     // first entry in the expanded task list (nblocks * 256 entries)
};

struct TaskSet {
    double threshold = -1;           // tasks with threshold <= Q_p Q_q < upper
    double upper = 0;
    bool sliced = false;             // J/K tasks: heavily contracted quartets split into primitive slices
    std::vector<QuartetClassTasks> classes;
    DevBuf<int64_t> d_prefix;
    DevBuf<int> d_koff, d_kbeg;
    DevBuf<hfgpu::TaskSeg> d_seg;
    DevBuf<int2> d_tasks;
};

struct PairClassHost {
    std::vector<int> sa, sb, prim0, nprim;
    std::vector<double> Q;
    std::vector<int> group;   // contraction groups: pairs [group[g], group[g+1]) share nprim
    DevBuf<int> d_sa, d_sb, d_prim0, d_nprim, d_group;
    DevBuf<hfgpu::PairRec> d_rec;
    DevBuf<double> d_Q;
    int npairs() const { return (int)sa.size(); }
};

struct KernelStat {
    int cls = 0;
    int fp32 = 0;             // quartets evaluated in FP32 (mixed-precision tier)
    int64_t launches = 0, quartets = 0, prim_quartets = 0;
    double eris = 0, flops = 0, ms = 0;
};

struct ScfWorkspace;   // the SCF's device matrices and events, kept between hf_scf calls (scf.hip)
struct ScfRun;         // an SCF driven step by step by the caller (hf_scf_begin / hf_scf_step / hf_scf_end)

struct hf_ctx {
    hf_options opt;
    std::string err;
    hipStream_t stream = nullptr;
    rocblas_handle blas = nullptr;
    bool have_mol = false;
    hfhost::Molecule mol;
    // device molecule data
    DevBuf<double> d_shell_xyz;
    DevBuf<int> d_shell_bf;
    DevBuf<hfgpu::PrimPair> d_pp;
    DevBuf<double> d_boys;
    DevBuf<float> d_boys32;
    std::vector<PairClassHost> pcls;
    TaskSet tasks_all;                     // unscreened (packed ERIs)
    std::map<int, TaskSet> tasks_by_level; // screened, by threshold level k (10^(k/4))
    std::map<int64_t, TaskSet> tasks_mixed_hi, tasks_mixed_lo;   // mixed precision: FP64 part, FP32 band
    TaskSet tasks_incore;                  // screened, whole quartets (in-core ERI matrix)
    // in-core ERIs: symmetric matrix over canonical function pairs, FP64 and its FP32 copy
    DevBuf<double> d_G;
    DevBuf<float> d_G32;
    DevBuf<double> d_Dw;                   // packed weighted total density of an in-core J/K build
    // per padded pair position: the Schwarz factor of its shell pair, rounded up to FP32 (0 for padding), and the
    // largest; a store element (P, Q) whose product of the two is below the threshold was screened out (zero)
    DevBuf<float> d_qf;
    double qf_max = 0;
    int64_t qf_live = 0;                   // store rows the contraction reads (qf_P qf_max at or above the threshold)
    bool qf_valid = false;
    hipStream_t eri_stream = nullptr;      // the SCF evaluates the in-core store here, beside its own stream
    double* hpin = nullptr;                // kPinnedDoubles of pinned host memory for small read-backs
    // pinned staging of host-to-device uploads (hf_upload): bytes [0, stage_used) belong to copies queued on the
    // context's stream
    char* stage = nullptr;
    size_t stage_cap = 0, stage_used = 0;
    // in-core contraction statistics ([0] FP64 store, [1] FP32 store): launches, elements read
    // (eris), algorithmic flops, and with profile = 1 HIP-event time
    KernelStat incore_stat[2];
    hipEvent_t incore_ev[2] = {nullptr, nullptr};
    int incore_pending = -1;               // stat index whose last launch's events are not yet read
    // J/K scratch
    DevBuf<double> d_Jp, d_Kp, d_D, d_J, d_K;
    DevBuf<long long> d_fx;          // deterministic mode: fixed-point J'/K' (direct) or K (in-core) accumulators
    // one-electron kernel inputs: per-shell L/bf/prim0/nprim, and alpha|coef|Z|atom xyz
    DevBuf<int> d_oe_L;
    DevBuf<double> d_oe_prim;
    size_t oe_nprim = 0;
    DevBuf<int2> d_oe_pairs;               // shell pairs by one-electron class (La, Lb), La >= Lb
    std::vector<int> oe_class_off;
    bool oe_valid = false;                 // the one-electron inputs above describe the current molecule
    DevBuf<double> d_Qdiag;                // Schwarz diagonals of one pair class (molecule set-up)
    // Device memory is kept between calls and molecules (buffers only grow): hipFree synchronises the
    // whole device, which would stall the other contexts of a server at every request.
    ScfWorkspace* scf_ws = nullptr;
    ScfRun* scf_run = nullptr;
    std::vector<KernelStat> stats;
    std::vector<hipEvent_t> ev;
    std::vector<int> ev_pending;         // stat slots whose last launch's events are not yet read
    // side streams: the class kernels of one pass run concurrently, forked from and joined
    // back to the caller's stream with events
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> side_ev;   // [0] fork, [1..] join per side stream
    // A device entry point called on a stream other than the context's own: the call's work is ordered after
    // the context stream's queued work and before any later work on it (they share the context's scratch)
    hipEvent_t order_ev[2] = {nullptr, nullptr};
    double last_jk_ms = 0;

    int fail(int code, const std::string& msg) {
        err = msg;
        return code;
    }
};

// Implemented in context.hip
// Small dense FP64 products on the matrix cores (csrc/dense.hip): C = alpha op(A) op(B) + beta C,
// column-major as BLAS.
int hf_dgemm_small(hf_ctx* ctx, hipStream_t st, bool ta, bool tb, int m, int n, int k, double alpha, const double* A,
                   int lda, const double* B, int ldb, double beta, double* C, int ldc);
constexpr int kSmallDense = 384;   // the matrix-core product kernel and the batched SCF kernels serve n <= this

// J/K of the densities d_D (ndens x n x n) on stream st, screened at Schwarz product thr (0: the
// context's threshold); accumulate adds to d_J/d_K instead of overwriting them; fp32 evaluates
// the quartets whose bound is below the context's fp32_threshold in FP32.
int hf_jk_device_impl(hf_ctx* ctx, int ndens, const double* d_D, double* d_J, double* d_K, hipStream_t st, double thr,
                      bool accumulate, bool fp32);
// In-core ERIs: G[P * np + Q] = (ij|kl) for canonical pairs P = i(i+1)/2 + j (i >= j) and
// Q = k(k+1)/2 + l (k >= l), np = n(n+1)/2; quartets below the context's Schwarz threshold are 0.
// len must be at least np * np.
int hf_eri_pairs_impl(hf_ctx* ctx, double* d_G, uint64_t len, hipStream_t st);
// J/K from in-core ERIs (G FP64, or FP32 when g32 is set) on stream st: J of the summed densities,
// K per density, both complete (no symmetrisation needed); accumulate adds to d_J/d_K.
// own_store: G is this context's store, evaluated at its Schwarz threshold, so the loads of elements that screening
// zeroed can be left out (same results, fewer bytes); false for a store the caller brings.
int hf_incore_jk_impl(hf_ctx* ctx, int ndens, const double* G, const float* g32, const double* d_D, double* d_J,
                      double* d_K, hipStream_t st, bool accumulate, bool own_store);
// One molecule's in-core build in hf_incore_jk_batch (arguments as hf_incore_jk_impl).
struct IncoreJob {
    hf_ctx* ctx;
    int ndens;
    const double* G;
    const float* g32;
    const double* D;
    double* J;
    double* K;
    bool accumulate;
    bool own_store;
};
// The in-core builds of several molecules on stream st: one launch packs every weighted density and clears the
// outputs, then one contraction launch per kind of build (store precision, densities, tile shape, column chunks)
// covers all the molecules of that kind.  The same sums per molecule as hf_incore_jk_impl (FP64 atomics aside).
int hf_incore_jk_batch(const std::vector<IncoreJob>& jobs, hipStream_t st);
// Takes the device address of this device's allocation-free scratch array (once per device; hf_create).
int hf_incore_scratch_init(hf_ctx* ctx);
// Ends an SCF session the caller left open (hf_destroy, hf_scf_begin).
void hf_scf_session_free(hf_ctx* ctx);
// Frees the SCF workspace (hf_destroy).
void hf_scf_workspace_free(hf_ctx* ctx);
// Device bytes the context holds between calls: SCF workspace, in-core stores, J/K scratch.
double hf_ctx_held_bytes(const hf_ctx* ctx);
// Frees the context's in-core stores when together they hold more than keep_bytes (a server calls
// this after each request, so idle slots do not hold large stores outside its admission budget).
void hf_release_incore(hf_ctx* ctx, size_t keep_bytes);
// Largest basis the in-core contraction kernel takes (its column registers cover 64 * 8 functions).
constexpr int kIncoreMaxN = 512;
// Tasks with threshold <= Q_p Q_q (< upper when upper > 0).
int hf_build_tasks(hf_ctx* ctx, double threshold, TaskSet& ts, double upper = 0.0, bool sliced = false);
// Implemented in one_electron.hip: S, T, V (n x n, device) on stream st; with dM (3 n x n) also
// the dipole moment integrals about the origin (x, y, z).
int hf_one_electron_device(hf_ctx* ctx, double* dS, double* dT, double* dV, hipStream_t st, double* dM = nullptr);
// Implemented in post.hip: MP2 sum over the nocc occupied orbitals of C (column-major, orbital p at
// C + p n) with energies eps, from the in-core store G; same_spin: the reference's UHF per-spin sum.
int hf_mp2_impl(hf_ctx* ctx, hipStream_t st, int nocc, const double* C, const double* eps, const double* G,
                bool same_spin, double& e2);
// Mulliken charges (natoms) and dipole (3) from the device densities d_D (ndens x n x n).
int hf_properties_impl(hf_ctx* ctx, int ndens, const double* d_D, double* charges, double* dipole, hipStream_t st);

// Orders work on a caller's stream st against the context's stream (both directions; no host wait): enter before
// the call's first launch, leave after its last.  Nothing to do when st is the context's stream.
int hf_stream_enter(hf_ctx* ctx, hipStream_t st);
int hf_stream_leave(hf_ctx* ctx, hipStream_t st);

// Host-to-device upload of bytes from src to dst, queued on the context's stream without waiting for it: the
// bytes are staged in the context's pinned buffer (the copy engine reads it directly; a synchronous hipMemcpy
// from pageable memory would wait for the copy, behind whatever the device is running).  src may be reused at
// once; the staging space is recycled after a wait for the stream when it runs out.
int hf_upload(hf_ctx* ctx, void* dst, const void* src, size_t bytes);
// Pinned host space for a device-to-host copy on the context's stream (valid until the next hf_upload or
// hf_stage_host after a wait for that copy).
void* hf_stage_host(hf_ctx* ctx, size_t bytes);

#define HF_HIP_CHECK(ctx, call)                                                                         \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return (ctx)->fail(HF_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));            \
    } while (0)

// Small device-to-host read-backs of the SCF loop (reduction results, stop flags): through the
// context's pinned buffer, which the copy engine writes directly (pageable memory would take a
// staging copy on every read-back).  count <= kPinnedDoubles.
constexpr int kPinnedDoubles = 64;
inline int hf_readback(hf_ctx* ctx, hipStream_t st, double* dst, const double* src, int count) {
    if (!ctx->hpin || count > kPinnedDoubles) {
        HF_HIP_CHECK(ctx, hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToHost, st));
        HF_HIP_CHECK(ctx, hipStreamSynchronize(st));
        return HF_OK;
    }
    HF_HIP_CHECK(ctx, hipMemcpyAsync(ctx->hpin, src, count * sizeof(double), hipMemcpyDeviceToHost, st));
    HF_HIP_CHECK(ctx, hipStreamSynchronize(st));
    for (int i = 0; i < count; ++i) dst[i] = ctx->hpin[i];
    return HF_OK;
}
