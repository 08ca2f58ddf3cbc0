// SCF session state shared by the one-molecule driver (scf.hip) and the lockstep batch of molecules
// (batch.hip): device matrices of one SCF, its DIIS history, the reference's iteration control as a
// state machine, and the step phases that run over a list of molecules.
#pragma once

#include <chrono>
#include <deque>
#include <vector>

#include "hf_internal.h"

constexpr int kSp2MaxSteps = 100;     // purification steps before a density step falls back to diagonalisation
constexpr int kSp2LocalMax = 96;      // n up to which all SP2 steps run in one workgroup (X in LDS: n^2 doubles)
constexpr int kWorkspaceMinN = 256;   // SCF workspace matrices are allocated for at least this many functions

// DIIS history of one spin (HartreeFock/DIIS.h:9-96): device slots for F and e, host inner products.
struct Diis {
    static constexpr int kKeep = 6, kFirst = 5;
    double* F = nullptr;          // kKeep * nn each (SCF workspace)
    double* E = nullptr;
    std::deque<int> order;        // slot indices, oldest first
    double B[kKeep][kKeep] = {};  // B[slot][slot]
    size_t nn = 0;
    int next_slot() const {
        if ((int)order.size() < kKeep) return (int)order.size();
        return order.front();
    }
};

struct ScfState {
    int n = 0;
    size_t nn = 0;
    DevBuf<double> S, h, X, F0, Fa, Fb, J, K, D, Dn, Dprev, dD, C, eps, T1, T2, T3, err, work, red, redpart, sp2, sp2W;
    DevBuf<rocblas_int> info;
};

// Device scratch of the step phases for a batch of `cap` spin jobs (a job = one molecule and one spin):
// reduction partials and results, SP2 step histories and partial traces, and a pinned host buffer
// for the read-backs.  Grows, never shrinks.
struct BatchScratch {
    int cap = 0;
    DevBuf<double> part, out, hist, spart, res;
    double* hpin = nullptr;
    hipStream_t side = nullptr;                 // single-workgroup purifications beside the stepped ones
    hipEvent_t ev[2] = {nullptr, nullptr};      // fork and join of the side stream
    ~BatchScratch();
    int reserve(hf_ctx* ctx, int jobs);
    int side_stream(hf_ctx* ctx);
};

// Device matrices and events of hf_scf, kept in the context between calls (buffers only grow), so a
// context that solves one molecule after another (a serving slot) allocates and frees nothing per SCF.
struct ScfWorkspace {
    ScfState st;
    DevBuf<double> diisF[2], diisE[2];
    BatchScratch bs;             // step-phase scratch of this context's own (one-molecule) steps
    hipEvent_t jk_ev[2] = {nullptr, nullptr};
    hipEvent_t eri_ev[2] = {nullptr, nullptr};
    hipEvent_t ready_ev = nullptr;   // end of Init on the context's stream (a batch's stream waits for it)
    ~ScfWorkspace();
};

// The J/K build of one molecule's step (ScfRun::jk_plan decides it, the step's first phase runs it).
struct JkPlan {
    enum Kind { kNone, kIncore, kDirect } kind = kNone;
    bool wait_store = false;      // the in-core store (evaluated on the context's ERI stream) must be complete first
    bool sub = false;             // dD = D - Dprev before the build (incremental build)
    bool copy_prev = false;       // Dprev <- D after the build (incremental builds follow)
    bool accumulate = false;      // add to the previous J/K
    const double* src = nullptr;  // the density contracted: D, or dD for an incremental build
    const float* g32 = nullptr;   // in-core: the FP32 copy of the store
    double thr = 0;               // direct: Schwarz product threshold (0: the context's)
    bool fp32 = false;            // direct: the mixed tier's FP32 quartets
    double q0 = 0;                // quartets evaluated before the build
};

// One SCF (HartreeFockAlgorithm::Init + Calculate, HartreeFock/HartreeFockAlgorithm.cpp:21-192) as a session
// object: init() is Init, a step is one RHF/UHF Step (run by scf_step_phases over one molecule or a batch),
// after_step() the reference's iteration control of Calculate and finish() the orbital energies, <S^2>, MP2
// and the result.  hf_scf runs them in turn; hf_scf_begin / hf_scf_step / hf_scf_end let a caller drive the
// steps itself (HartreeFockAlgorithm::Step); a batch (hf_scf_batch, the server) steps several sessions in
// lockstep.
struct ScfRun {
    hf_ctx* ctx;
    hf_scf_params P;
    const hf_scf_params* p;
    std::chrono::steady_clock::time_point t_start;
    hipStream_t s = nullptr;     // the context's stream (Init, finish)
    bool uhf = false, mixed_tier = false, incore = false, store_pending = false, field = false;
    int n = 0, nsp = 1, ne = 0, occ[2] = {0, 0};
    size_t nn = 0;
    double Enuc = 0, eri_ms = 0;
    hipEvent_t* eri_ev = nullptr;
    hipEvent_t* jk_ev = nullptr;
    Diis diis[2];
    double etolD = 0, etolN = 0, dtol = 0, itol = 0;
    double energy = 0, homo = 0, lastErr = 0, jk_ms = 0;
    int diag_steps = 0, sp2_steps[2] = {0, 0}, sp2_total = 0;
    bool last_diag[2] = {false, false};
    // incremental Fock build state: J, K hold the J/K of Dprev
    int full_builds = 0, incr_builds = 0, since_full = 0;
    bool force_full = false, last_incremental = false;
    double quartets = 0, last_dmax = 0, last_quartets = 0;
    bool incore_mixed = false, incremental = false;
    int first_full = 1, rebuild_every = 8;
    static constexpr double kIncoreFp32MaxDelta = 1e-3;   // largest max|D - Dprev| contracted with the FP32 store
    bool jk_pending = false;
    // mixed-precision tier: quartets below the context's fp32_threshold in FP32 until fp64_phase
    bool mixed = false, fp64_phase = true, last_fp32 = false;
    int final_fp64 = 3, fp64_steps = 0, fp32_builds = 0;
    bool useDIIS = true;
    std::vector<double> eps_h;
    // iteration state of the iteration control (or of a caller's steps), read by finish()
    int iter = 0;
    bool converged = false, cancelled = false;
    double cur = 0;
    bool external = false;   // steps driven by the caller (hf_scf_step): no convergence verdict of its own
    bool orbitals = true;    // finish(): orbital energies (HOMO), <S^2> and MP2; a server's results need none

    // ---- one step's transient state, written by the step phases ----
    int step_iter = 0;           // iteration the current step runs
    int status = HF_OK;          // a failure of this molecule's step (the batch carries on without it)
    bool diis_on = false, diis_used = false;
    bool sp2_ok[2] = {false, false};
    const double* sp2_P[2] = {nullptr, nullptr};
    double rms = 0;              // ||D_new - D||_F of the step
    // max|D - Dprev| of the next incremental J/K build, known from the step's energy phase:
    // D_new - D_old = (Dn - D_old) with DIIS, alpha (Dn - D_old) with mixing (-1: not known)
    double dmax_next = -1;

    // ---- iteration control of Calculate (HartreeFock/HartreeFockAlgorithm.cpp:105-192) ----
    enum Phase { kPhDiis, kPhAfterDiis, kPhPlain, kPhDone };
    Phase ph = kPhDiis;
    int ph_i = 0;
    double prev = 0;
    bool trace = false;

    ScfRun(hf_ctx* c, const hf_scf_params& params) : ctx(c), P(params), p(&P) {}
    double quartet_count() const {
        double q = 0;
        for (const auto& ks : ctx->stats) q += (double)ks.quartets;
        return q;
    }
    int init();
    // Iteration control: begin() before the first step; next_iter() is the iteration the next step runs;
    // after_step() digests a finished step (energy, rms) and returns an error for a non-finite SCF.
    void begin();
    int next_iter() const { return ph == kPhAfterDiis ? iter + ph_i : iter; }
    bool done() const { return ph == kPhDone; }
    int after_step();
    int calculate();
    int finish(hf_scf_result* r, double* D_out);
    // The J/K build of the current density for the coming Fock matrix (full, incremental, FP64 or FP32 as the
    // SCF's state says), with its bookkeeping; jk_done after it ran.
    void jk_plan(JkPlan& pl);
    void jk_done(const JkPlan& pl);
    // The one-molecule step (hf_scf_step, calculate()): the step phases over this molecule alone.
    int step(int iter, double& rms);
};

// One SCF step (HartreeFock/RestrictedHartreeFock.cpp:61-166, UnrestrictedHartreeFock.cpp:87-191) of every
// molecule in ms, each at its own step_iter, on stream st: J/K and Fock matrices, DIIS, the new density and
// the energy.  Every phase is one launch over all the molecules (J/K builds are per molecule); a molecule
// whose step fails gets its status set and is left out of the remaining phases.  Returns an error only for
// a failure of the stream itself (then every molecule's step is void).
int scf_step_phases(const std::vector<ScfRun*>& ms, hipStream_t st, BatchScratch& bs);

// Dense helpers of the SCF on an explicit stream: C = alpha op(A) op(B) + beta C (column-major; the
// matrix-core kernel for n <= kSmallDense, else rocBLAS), and the symmetric eigenproblem in place.
int scf_gemm(hf_ctx* ctx, hipStream_t st, bool ta, bool tb, int m, int n, int k, double alpha, const double* A,
             const double* B, double beta, double* C);
int scf_syevd(hf_ctx* ctx, ScfState& ss, hipStream_t st, double* A, double* w);
// SP2 projector through library products (n > kSmallDense), on stream st.
int scf_sp2_projector(hf_ctx* ctx, ScfState& ss, hipStream_t st, const double* Fp, int nocc, double* P, double* W,
                      bool& ok);
// The single-workgroup SP2 kernel's dynamic-LDS limit is set (once per process); false if that failed.
bool sp2_local_ready();

// Batched dense kernels (batch.hip).  A job names one molecule's matrices; up to kBatchJobs jobs per launch
// (more are split over launches).
constexpr int kBatchJobs = 32;
