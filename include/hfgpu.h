/*
 * hfgpu.h — C ABI of the MI355X (gfx950) two-electron integral / J-K Fock-build
 * engine and the device-resident SCF built on it.
 *
 * The reference (aromanro/HartreeFock) has no FFI: its hot path is reached through
 * in-process C++ classes.  This header is the thin C layer those classes sit on;
 * each entry point names the reference member it replaces.  Conventions:
 *   - every function returns HF_OK (0) or a negative HF_E_* code; hf_last_error()
 *     gives the message; no C++ exception crosses this boundary;
 *   - matrices are n x n doubles, row-major (all exchanged matrices are symmetric);
 *   - coordinates are bohr, as in the reference's Molecule;
 *   - basis functions are ordered atom -> shell (file order) -> Cartesian component
 *     (xx,xy,xz,yy,yz,zz), the order of HartreeFock/Molecule.cpp:41-72;
 *   - one context per host thread; a context owns all device memory it uses.
 */
#ifndef HFGPU_H
#define HFGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HF_OK 0
#define HF_E_ARG (-1)
#define HF_E_OOM (-2)
#define HF_E_HIP (-3)
#define HF_E_NOT_CONVERGED (-4)
#define HF_E_CANCELLED (-5)
#define HF_E_IO (-6)
#define HF_E_STATE (-7)
#define HF_E_UNSUPPORTED (-8)
#define HF_E_NUMERIC (-9)      /* a non-finite energy or density: the request failed, nothing else is touched */

#define HF_PREC_FP64 0
#define HF_PREC_MIXED 1

typedef struct hf_ctx hf_ctx;

typedef struct hf_options {
    int device;                /* HIP device ordinal */
    double schwarz_threshold;  /* direct J/K: skip quartets with Q_ab*Q_cd below this (0 = none) */
    double prim_threshold;     /* drop primitive pairs with |c_a c_b N_a N_b e^{-ab/p AB^2}| below this */
    int exact_transfer;        /* 0: the reference's ERI convention (see hf_eri_packed); 1: exact ERIs */
    int precision;             /* HF_PREC_FP64 or HF_PREC_MIXED */
    double fp32_threshold;     /* mixed: quartets whose bound is below this run in FP32 */
    int profile;               /* 1: time class kernels with HIP events (hf_kernel_stats) */
    int profile_class;         /* with profile = 1: time only the class la*1000 + lb*100 + lc*10 + ld
                                  (-1, the default: every class; -2: no class kernel, only the in-core
                                  contraction); each timed launch is bracketed by two events, which cost
                                  the build a few per cent */
    int incore_shape;          /* in-core contraction tile shape: -1 (the default) lets the autotuner time
                                  the candidates once per process and size; 0..7 forces one, for tests of
                                  every shape and for measurements.  (columns per lane, rows per group,
                                  two register buffers in turn, store rows per block): 0 (2,2,no,1),
                                  1 (2,2,yes,1), 2 (2,1,yes,1), 3 (1,2,no,1), 4 (1,2,yes,1), 5 (2,1,no,3),
                                  6 (2,1,yes,3), 7 (2,2,no,3).  Shapes 3 and 4 need n <= 512 with one
                                  density and n <= 256 with two, shapes 5..7 n <= 256 with two densities
                                  (else shape 0 runs); other values fail with HF_E_ARG */
    int deterministic;         /* 1: J/K builds and the nuclear-attraction matrix accumulate in a fixed-point
                                  form whose sums do not depend on the order of the GPU's additions (two
                                  64-bit integers per element, 2^-80 resolution), so repeated runs give
                                  bit-identical energies; the FP64 atomics' order varies from run to run.
                                  Slower (two integer atomics per contribution, no LDS staging of exchange
                                  rows); not available to hf_build_jk_incore_device_noalloc */
} hf_options;

void hf_default_options(hf_options* opt);

/* Context lifetime.  hf_create replaces constructing a HartreeFockAlgorithm (its
 * IntegralsRepository member, HartreeFock/HartreeFockAlgorithm.h:51). */
int hf_create(const hf_options* opt, hf_ctx** out);
int hf_destroy(hf_ctx* ctx);
const char* hf_last_error(const hf_ctx* ctx);

/* Shells with SP shells already split into S and P (same exponents).  coef holds the
 * contraction coefficient times the radial normalisation (2a/pi)^{3/4} (4a)^{L/2};
 * the per-component factor 1/sqrt((2l-1)!!(2m-1)!!(2n-1)!!) is applied by the library
 * (together these are the reference's per-primitive factor, HartreeFock/GaussianOrbital.cpp:25-30). */
typedef struct hf_basis_soa {
    int nshell;
    int nbf;
    const int* L;              /* nshell */
    const int* nprim;          /* nshell */
    const int* prim_offset;    /* nshell: first primitive in alpha/coef */
    const int* bf_offset;      /* nshell: first basis function */
    const int* atom;           /* nshell: atom index */
    const double* xyz;         /* 3*nshell, bohr */
    const double* alpha;       /* exponents */
    const double* coef;        /* coefficient x radial normalisation */
    int natoms;
    const int* Z;              /* natoms: nuclear charges */
    const double* atom_xyz;    /* 3*natoms, bohr */
} hf_basis_soa;

/* Replaces IntegralsRepository::Reset(Molecule*) (HartreeFock/IntegralsRepository.cpp:19-27):
 * installs a molecule, uploads shells, primitive pairs and the Schwarz table. */
int hf_set_basis(hf_ctx* ctx, const hf_basis_soa* basis);
/* Same, parsing an NWChem basis file as Chemistry::Basis::Load (HartreeFock/Basis.cpp:15-158)
 * and placing atoms as Molecule::LoadXYZ (HartreeFock/Molecule.cpp:143-199). */
int hf_set_molecule(hf_ctx* ctx, const char* basis_file, int natoms, const int* Z, const double* xyz);
int hf_nbf(const hf_ctx* ctx);
int hf_nshell(const hf_ctx* ctx);
/* Electrons as Molecule::Init (HartreeFock/Molecule.cpp:127-140): alpha = N/2, beta = N - alpha, or the
 * counts given to hf_set_electrons; returns N. */
int hf_electrons(const hf_ctx* ctx, int* alpha, int* beta);
double hf_nuclear_repulsion(const hf_ctx* ctx);   /* Molecule::NuclearRepulsionEnergy */
/* Molecule::alphaElectrons / betaElectrons (HartreeFock/Molecule.h:16-17) of the installed molecule, as
 * RestrictedHartreeFock::Init (N = alpha + beta, N/2 doubly occupied; HartreeFock/RestrictedHartreeFock.cpp:23-31)
 * and UnrestrictedHartreeFock::Init (alpha and beta occupied; HartreeFock/UnrestrictedHartreeFock.cpp:29-58) read
 * them: ions and high-spin states.  Both 0 (the default): the neutral count, split as Molecule::Init.  Reset by
 * hf_set_molecule / hf_set_basis. */
int hf_set_electrons(hf_ctx* ctx, int alpha, int beta);
/* Molecule::ElectricField (HartreeFock/Molecule.h:19), atomic units: the SCF's core Hamiltonian becomes
 * h = T + V + E . <mu|r|nu> with the moment integrals about the origin (HartreeFock/HartreeFockAlgorithm.cpp:38).
 * Reset by hf_set_molecule / hf_set_basis. */
int hf_set_electric_field(hf_ctx* ctx, const double* field3);
/* Molecule::NuclearElectricFieldEnergy (HartreeFock/Molecule.cpp:86-94), -sum_A Z_A R_A . E; as in the
 * reference it is not part of the SCF's total energy (HartreeFock/RestrictedHartreeFock.cpp:248). */
double hf_nuclear_electric_field_energy(const hf_ctx* ctx);

/* Packed unique ERIs at index GetElectronElectronIndex(i,j,k,l)
 * (HartreeFock/IntegralsRepository.h:151-162).  Replaces
 * IntegralsRepository::CalculateElectronElectronIntegrals (HartreeFock/IntegralsRepository.cpp:566-584)
 * followed by reads through getElectronElectron(int,int,int,int) (:138-141).
 * With exact_transfer = 0 the (dd|dd) values follow the reference's electron-transfer
 * range (HartreeFock/GaussianTwoElectrons.cpp:152), which leaves the rows |e|=2 at
 * |f|=4 at zero; exact_transfer = 1 gives the exact integrals. */
uint64_t hf_packed_index(uint32_t i, uint32_t j, uint32_t k, uint32_t l);
uint64_t hf_packed_len(int nbf);           /* GetElectronElectronIndex(n,n,n,n)+1, as the reference sizes it */
int hf_eri_packed(hf_ctx* ctx, double* out, uint64_t len);                 /* host buffer */
int hf_eri_packed_device(hf_ctx* ctx, double* d_out, uint64_t len, void* stream);

/* Direct J/K build, replacing the N^4 loops of RestrictedHartreeFock::InitFockMatrix
 * (HartreeFock/RestrictedHartreeFock.cpp:189-202) and UnrestrictedHartreeFock::InitFockMatrices
 * (HartreeFock/UnrestrictedHartreeFock.cpp:228-238):
 *   J  = J[D_0 + ... + D_{ndens-1}],  J_ij = sum_kl (ij|kl) D_kl
 *   K_d = K[D_d],                      K_ij = sum_kl (il|kj) D_kl
 * ndens is 1 (RHF) or 2 (UHF: alpha and beta in one pass).  D is ndens*n*n, J n*n,
 * K ndens*n*n.  Quartets with Q_ab*Q_cd < schwarz_threshold are skipped. */
int hf_build_jk(hf_ctx* ctx, int ndens, const double* D, double* J, double* K);             /* host */
int hf_build_jk_device(hf_ctx* ctx, int ndens, const double* d_D, double* d_J, double* d_K,
                       void* stream);                                                       /* device */

/* In-core J/K: the reference's scheme of storing every ERI once and contracting the store with the
 * density in each Fock build (HartreeFock/IntegralsRepository.cpp:566-584, then
 * HartreeFock/RestrictedHartreeFock.cpp:189-202 and HartreeFock/UnrestrictedHartreeFock.cpp:228-238),
 * with the store in HBM.  The store is the symmetric matrix over canonical function pairs,
 * G[P*np + Q] = (ij|kl) with P = pos(i, j) (i >= j), Q = pos(k, l) (k >= l), where the pair triangle is
 * laid out row by row with each row padded to an even length: pos(k, l) = base(k) + l, base(k) =
 * k(k+2)/2 for even k and (k+1)^2/2 for odd k, np = base(n) (padding entries are 0).
 * hf_pair_matrix_len(n) = np*np doubles: 2.9 GB for n = 194, 137 GB for n = 512.  Screened at the
 * context's Schwarz threshold.  J and K as hf_build_jk; n <= 512. */
uint64_t hf_pair_matrix_len(int nbf);
int hf_eri_pairs_device(hf_ctx* ctx, double* d_G, uint64_t len, void* stream);
int hf_build_jk_incore_device(hf_ctx* ctx, int ndens, const double* d_G, const double* d_D, double* d_J,
                              double* d_K, void* stream);
/* Host convenience: evaluates the store in the context's memory, then contracts. */
int hf_build_jk_incore(hf_ctx* ctx, int ndens, const double* D, double* J, double* K);
/* hf_build_jk_incore_device without any allocation: no counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops that forbid hipMalloc).  The packed
 * weighted density it needs lives in a static device array of 131584 doubles (1 MiB) per GPU, so n <= 512
 * (larger n is rejected with HF_E_ARG); the tile shape is the autotuner's if a build of this kind has already
 * run in the process (else the default shape), and no profiling events are used.  Every call on one GPU shares
 * that array: calls that overlap on one GPU (other streams, other contexts or threads) race and must be
 * serialised by the caller.  Same results as hf_build_jk_incore_device. */
int hf_build_jk_incore_device_noalloc(hf_ctx* ctx, int ndens, const double* d_G, const double* d_D, double* d_J,
                                      double* d_K, void* stream);

/* One-electron matrices S, T, V (HartreeFock/QuantumMatrix.cpp:34-155). Host buffers, n*n. */
int hf_one_electron(hf_ctx* ctx, double* S, double* T, double* V);
/* The core Hamiltonian the SCF starts from, h = T + V + E . M (the public HartreeFockAlgorithm::h,
 * HartreeFock/HartreeFockAlgorithm.h:47, set at :38), n*n host buffer. */
int hf_core_hamiltonian(hf_ctx* ctx, double* h);
/* Boys functions F_0..F_mmax(T) (mmax <= 32), as IntegralsRepository::getBoysFunctions / BoysFunctions
 * (HartreeFock/IntegralsRepository.cpp:326-342, HartreeFock/BoysFunctions.cpp:10-25); computed on the host. */
int hf_boys(int mmax, double T, double* F);

/* SCF, replacing RestrictedHartreeFock / UnrestrictedHartreeFock Init() + Calculate()
 * (HartreeFock/HartreeFockAlgorithm.cpp:21-192).  Field meanings and defaults follow the
 * reference library (HartreeFock/HartreeFockAlgorithm.cpp:10-11, UnrestrictedHartreeFock.cpp:9). */
typedef struct hf_scf_params {
    int unrestricted;              /* 0 RHF, 1 UHF */
    int max_iterations;            /* 3000 */
    double alpha;                  /* 0.75: density mixing when DIIS is not used */
    double init_guess;             /* 0.75: iteration-0 F = init_guess*S_ij*(h_ii+h_jj)/2; 0 -> F = h */
    int use_diis;                  /* 1 */
    int max_diis_iterations;       /* 1000 */
    int normal_iter_after_diis;    /* 500 */
    int add_asymmetry;             /* UHF: 1 */
    double asymmetry;              /* UHF: 0.1 */
    double energy_tolerance;       /* 0: the reference's (1e-11 with DIIS, 1e-13 without) */
    double density_tolerance;      /* 0: the reference's 1e-10 (Frobenius norm of the change) */
    double diis_tolerance;         /* 0: the reference's 1e-10 */
    volatile int* terminate;       /* optional cancellation flag (another thread sets it non-zero), read
                                      with an atomic load once per iteration; the SCF then returns
                                      HF_E_CANCELLED with converged = 0 */
    int density_method;            /* 0: spectral-projection purification (SP2), falling back to
                                      diagonalisation for a step it cannot resolve; 1: diagonalise
                                      every step (the reference's route).  Same fixed point. */
    int incremental_fock;          /* 1: between full builds, J/K of the density change only, screened
                                      at threshold / max|dD|; convergence is only accepted on a step
                                      with a full build.  0 (default): full J/K every step. */
    int full_rebuild_every;        /* incremental steps between full builds (default 8) */
    int final_fp64_iterations;     /* mixed precision (context option): all-FP64 steps required at the
                                      end before convergence is accepted (default 0: 3 with direct J/K,
                                      1 with the in-core store, whose FP64 steps are full FP64 builds) */
    int incore;                    /* 1: evaluate the ERIs once per SCF into HBM (hf_eri_pairs_device)
                                      and build J/K from the store when it fits; 0 (default): direct */
    double incore_max_bytes;       /* largest store the SCF allocates (0: half the free device memory;
                                      negative: none, the SCF stays direct) */
    int mp2;                       /* 1: after the SCF, the MP2 correlation energy as the reference's
                                      CalculateMp2Energy (HartreeFock/RestrictedHartreeFock.cpp:251-288,
                                      HartreeFock/UnrestrictedHartreeFock.cpp:295-375) from the in-core
                                      store (evaluated for it if the SCF ran direct; n <= 512) */
} hf_scf_params;

typedef struct hf_scf_result {
    double energy;                 /* total energy, Eh */
    double homo;                   /* HOMO energy, Eh */
    double s2;                     /* <S^2> for UHF (not in the reference; 0 for RHF) */
    int iterations;
    int converged;
    double jk_ms;                  /* time in J/K builds */
    double total_ms;               /* wall time of hf_scf */
    int fp32_builds;               /* mixed precision: J/K builds that used FP32 quartets */
    int diag_steps;                /* density steps that used diagonalisation */
    int full_builds;               /* J/K builds of the full density */
    int incremental_builds;        /* J/K builds of a density change */
    double quartets;               /* shell quartets evaluated over all J/K builds */
    int sp2_steps;                 /* purification steps over all density steps (their final steps) */
    int incore;                    /* 1: J/K came from the in-core ERI store */
    double eri_ms;                 /* in-core: time to evaluate the store */
    double mp2_energy;             /* with mp2 = 1: MP2 correlation energy, Eh */
} hf_scf_result;

void hf_scf_defaults(hf_scf_params* p);
/* D_out (optional): n*n (RHF total density) or 2*n*n (UHF alpha then beta). */
int hf_scf(hf_ctx* ctx, const hf_scf_params* p, hf_scf_result* r, double* D_out);

/* The same SCF driven one step at a time by the caller, for HartreeFockAlgorithm::Step(int)
 * (HartreeFock/HartreeFockAlgorithm.h:77; RestrictedHartreeFock::Step, HartreeFock/RestrictedHartreeFock.cpp:61-166,
 * UnrestrictedHartreeFock::Step, HartreeFock/UnrestrictedHartreeFock.cpp:87-191).  hf_scf_begin does Init (one-electron
 * integrals, S^-1/2, iteration-0 guess, the in-core store when asked); each hf_scf_step(iter) is one Step: the Fock
 * build of the current density (iter 0: the guess), DIIS while p->use_diis and iter < max_diis_iterations, the new
 * density and the energy, returning the energy, ||D_new - D||_F (Step's return value) and the DIIS error estimate;
 * hf_scf_end gives the result (orbital energies, <S^2>, the density; converged = 0: convergence is the caller's
 * call) and ends the session.  The steps are all FP64 (the mixed tier's switching belongs to hf_scf's own loop).
 * One session per context; hf_set_molecule / hf_set_basis / hf_destroy end an open one. */
int hf_scf_begin(hf_ctx* ctx, const hf_scf_params* p);
int hf_scf_step(hf_ctx* ctx, int iter, double* energy, double* rms, double* diis_error);
int hf_scf_end(hf_ctx* ctx, hf_scf_result* r, double* D_out);

/* Lockstep SCF of several molecules, each installed in its own context (count contexts, all different), with
 * its own parameters (params[i]): no counterpart in the reference's library, which runs one SCF per std::thread
 * (HartreeFock/HartreeFockDoc.cpp:379-401, HartreeFockThread.cpp:79-121); this replaces that loop of threads for
 * the molecules of one GPU.  Every molecule runs hf_scf's Init, steps and iteration control, but the steps of all
 * the molecules still iterating advance together on ctxs[0]'s stream, and each phase of an iteration (Fock
 * assembly, DIIS products and inner products, orthogonalisation, SP2 purification, density products, energy
 * reductions, density update) is one launch over all of them, with one host round trip per phase for the whole
 * batch; the J/K builds stay per molecule.  A molecule leaves the batch when its SCF ends (its result then comes
 * from its own context, as hf_scf's).  results[i] and status[i] (HF_OK, HF_E_NOT_CONVERGED, ...) per molecule;
 * returns HF_OK when every status is HF_OK, else the first non-OK status.  Each result equals the molecule's
 * hf_scf result to the FP64 atomics' rounding (in deterministic mode, bit for bit). */
int hf_scf_batch(hf_ctx* const* ctxs, int count, const hf_scf_params* params, hf_scf_result* results, int* status);

/* Mulliken charges (natoms) and dipole moment (3, au, about the origin) of the densities D (host,
 * ndens x n x n: the RHF total density, or UHF alpha then beta), replacing
 * RestrictedHartreeFock::CalculateAtomicCharge / GetMoment (HartreeFock/RestrictedHartreeFock.cpp:291-335)
 * and the UnrestrictedHartreeFock ones (HartreeFock/UnrestrictedHartreeFock.cpp:381-420).  Overlap and
 * dipole integrals are computed on the device.  charges or dipole may be null. */
int hf_properties(hf_ctx* ctx, int ndens, const double* D, double* charges, double* dipole);

/* Per-class ERI kernel statistics, summed over every launch since the last reset
 * (J/K builds, packed-ERI passes).  No counterpart in the reference (it has no timers);
 * used for ERI/s and roofline reporting. */
typedef struct hf_kernel_stat {
    int la, lb, lc, ld;            /* angular class (la lb | lc ld) */
    int64_t launches;
    int64_t quartets;              /* contracted shell quartets evaluated */
    int64_t prim_quartets;         /* primitive quartets evaluated */
    double eris;                   /* Cartesian ERIs evaluated (quartets x components) */
    double flops;                  /* algorithmic flops (generated-code operation count) */
    double ms;                     /* kernel time from HIP events on the launch stream (profile = 1), else 0 */
    int fp32;                      /* 1: quartets evaluated in FP32 (mixed precision) */
} hf_kernel_stat;
int hf_kernel_stats(const hf_ctx* ctx, hf_kernel_stat* out, int max);   /* returns count */
int hf_kernel_stats_reset(hf_ctx* ctx);

/* ---- Serving: continuous batching of SCF requests on one GPU ----
 * No counterpart in the reference's library (its GUI runs one SCF per std::thread,
 * HartreeFock/HartreeFockDoc.cpp:379-401).  A server owns `slots` contexts (stream, device memory, SCF
 * state kept on the device between requests), stepped together in lockstep (see lockstep) or each by its
 * own worker thread.  A request is
 * admitted to a free slot, first come first served, when its estimated HBM footprint fits in
 * hbm_budget together with the requests already running; requests that arrive while others run
 * are admitted as slots and memory free up, and the kernels of concurrent requests overlap on the
 * device.  Each result is the same as the molecule's SCF run alone (to the atomics' rounding). */
typedef struct hf_server hf_server;

typedef struct hf_serve_options {
    int slots;                     /* concurrent requests (contexts), default 32 */
    double hbm_budget;             /* bytes for admitted requests; 0 = 80% of the device's free memory */
    hf_options ctx;                /* context options of every slot (precision, thresholds, ...) */
    hf_scf_params scf;             /* SCF parameters of every request (unrestricted is per request) */
    int lockstep;                  /* k >= 1 (default 2): k batch threads, each with its own stream, step the
                                      admitted requests together: the SCF steps of a batch's requests advance
                                      together, each phase of an iteration one launch over all of them (as
                                      hf_scf_batch), and a prepared request joins the smallest batch at its next
                                      iteration boundary and leaves when its SCF ends (two batches fill each
                                      other's host round trips); prep_threads threads prepare newly admitted
                                      requests (molecule set-up, one-electron integrals, S^-1/2, in-core store)
                                      on their slots' streams.  0: one worker thread per slot, each request's SCF
                                      run alone (hf_scf) on its slot's stream, the slots' kernels overlapping */
    int prep_threads;              /* lockstep: preparation threads, default 8 */
} hf_serve_options;

typedef struct hf_serve_result {
    uint64_t id;                   /* from hf_serve_submit */
    uint64_t tag;                  /* caller's tag */
    int status;                    /* HF_OK, HF_E_NOT_CONVERGED or another error */
    int nbf;
    double energy;
    int iterations;
    int converged;
    double arrival_ms;             /* submit time, ms since the server was created */
    double admit_ms;               /* start of its work on a slot */
    double done_ms;                /* completion; latency = done_ms - arrival_ms */
    double footprint;              /* estimated device bytes while it ran */
} hf_serve_result;

void hf_serve_default_options(hf_serve_options* o);
int hf_serve_create(const hf_serve_options* o, hf_server** out);
/* Queue one molecule (atoms in bohr); never blocks on the GPU.  A request whose footprint
 * (hf_serve_footprint, plus its in-core store) exceeds the server's hbm_budget is rejected here with
 * HF_E_OOM: it could never be admitted. */
int hf_serve_submit(hf_server* s, const char* basis_file, int natoms, const int* Z, const double* xyz,
                    int unrestricted, uint64_t tag, uint64_t* id);
/* Up to max finished results, waiting up to wait_ms for the first one; returns the count. */
int hf_serve_poll(hf_server* s, hf_serve_result* out, int max, int wait_ms);
/* Kernel statistics (as hf_kernel_stats) of every request finished since the server started or since the
 * last call with reset = 1, summed over the slots; returns the count.  With profile = 1 in the slots'
 * context options the times come from HIP events on each slot's streams. */
int hf_serve_kernel_stats(hf_server* s, hf_kernel_stat* out, int max, int reset);
/* Requests submitted and not yet returned by hf_serve_poll. */
int hf_serve_pending(hf_server* s);
/* Estimated device bytes of an SCF on n basis functions (the admission rule); a server whose SCF
 * parameters ask for the in-core store adds hf_pair_matrix_len(n) doubles (x 1.5 in the mixed tier), and
 * the request's SCF takes the store only within that allowance.  A slot keeps its device memory (SCF
 * matrices, a store up to min(16 GiB, hbm_budget / slots)) for its next request; those bytes stay
 * charged to the budget while the slot is idle, and idle slots release their stores when the request at
 * the head of the queue would otherwise not fit. */
double hf_serve_footprint(int nbf, int unrestricted);
/* Waits for the running requests, drops the queued ones, frees everything. */
int hf_serve_destroy(hf_server* s);

#ifdef __cplusplus
}
#endif
#endif
