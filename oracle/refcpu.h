/*
 * refcpu — CPU oracle for the ERI / J-K / SCF hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, on the CPU, the algorithm of
 * the reference program (aromanro/HartreeFock) for the two-electron integrals,
 * the Fock build and the SCF loop around it.  It is the checker the GPU library
 * is compared against; nothing in the product links or calls it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg use it.
 *
 * Every entry point returns 0 on success and a negative number on error.
 * Matrices are row-major n*n doubles.  Coordinates are in bohr.
 */
#ifndef REFCPU_H
#define REFCPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ref_mol ref_mol;

/* Build a molecule: basis file (NWChem format) + atoms.  Follows
 * HartreeFock/Basis.cpp:15-158, HartreeFock/Molecule.cpp:41-140. */
int ref_mol_create(const char* basis_file, int natoms, const int* Z, const double* xyz,
                   ref_mol** out);
void ref_mol_destroy(ref_mol* m);
int ref_mol_nbf(const ref_mol* m);
int ref_mol_max_l(const ref_mol* m);
int ref_mol_electrons(const ref_mol* m, int* alpha, int* beta);
double ref_mol_nuclear_repulsion(const ref_mol* m);
/* Molecule::alphaElectrons / betaElectrons (HartreeFock/Molecule.h:16-17): ions, high-spin states. */
int ref_mol_set_electrons(ref_mol* m, int alpha, int beta);
/* Molecule::ElectricField (HartreeFock/Molecule.h:19): h = T + V + E . M in the SCF. */
int ref_mol_set_field(ref_mol* m, const double* E);
/* Per basis function: L, (l,m,n), atom index, shell index (reference IDs). */
int ref_mol_bf_info(const ref_mol* m, int* L, int* lmn, int* atom, int* shell);

/* The reference's electron-transfer step (HartreeFock/GaussianTwoElectrons.cpp:137-192)
 * fills bra rows e only for |e| >= |f|-1 at ket level |f|.  For (dd|dd) the rows
 * |e| = 2 at |f| = 4 are needed by the bra HRR but stay zero there, so two-centre-bra
 * (dd|dd) integrals differ from the exact ones.  Default (0) reproduces the reference;
 * 1 fills every row (exact integrals).  Process-global; not thread-safe. */
void ref_set_exact_transfer(int on);

/* Boys functions F_0..F_maxm(T) exactly as HartreeFock/BoysFunctions.cpp:10-25. */
int ref_boys(int maxm, double T, double* out);

/* One ERI between four basis functions (0-based IDs), as
 * IntegralsRepository::getElectronElectron(CGO*x4), HartreeFock/IntegralsRepository.cpp:372-433. */
double ref_eri(ref_mol* m, int i, int j, int k, int l);

/* Packed unique ERIs, HartreeFock/IntegralsRepository.cpp:566-584 with the index of
 * HartreeFock/IntegralsRepository.h:151-162.  len must be >= ref_packed_len(n). */
uint64_t ref_packed_index(uint32_t i, uint32_t j, uint32_t k, uint32_t l);
uint64_t ref_packed_len(int n);
int ref_eri_packed(ref_mol* m, double* out, uint64_t len);
/* Same, restricted to basis functions i in [i0, i1): the bounded CPU-baseline sample. */
int ref_eri_packed_rows(ref_mol* m, double* out, uint64_t len, int i0, int i1);

/* Every stride-th (i,j,k) triple of the packed loop with its whole l loop (timing sample). */
int ref_eri_packed_sample(ref_mol* m, double* out, uint64_t len, uint64_t stride, uint64_t offset,
                          uint64_t* count);
/* RHF G = sum_kl D_kl [(ij|kl) - 1/2 (il|kj)], the reference's N^4 loop, rows [i0, i1). */
int ref_rhf_g(int n, const double* eri, const double* D, double* G, int i0, int i1);

/* One-electron matrices S, T, V (HartreeFock/QuantumMatrix.cpp:34-155). */
int ref_one_electron(ref_mol* m, double* S, double* T, double* V);
/* Dipole moment integrals about the origin (HartreeFock/QuantumMatrix.cpp:61-90). */
int ref_moment(ref_mol* m, double* Mx, double* My, double* Mz);

/* J/K from a packed ERI store: J_ij = sum_kl (ij|kl) D_kl, K_ij = sum_kl (il|kj) D_kl,
 * the N^4 loop of HartreeFock/RestrictedHartreeFock.cpp:189-202. */
int ref_jk_from_packed(int n, const double* eri, const double* D, double* J, double* K);

typedef struct ref_scf_params {
    int unrestricted;       /* 0 RHF, 1 UHF */
    int max_iterations;     /* reference default 3000 */
    double alpha;           /* density mixing when DIIS is not used */
    double init_guess;      /* 0: F = h at iteration 0 */
    int use_diis;
    int max_diis_iterations;
    int normal_iter_after_diis;
    int add_asymmetry;      /* UHF only */
    double asymmetry;
} ref_scf_params;

typedef struct ref_scf_result {
    double energy;
    double homo;
    int iterations;
    int converged;
    double s2;              /* <S^2> (UHF), 0 for RHF */
} ref_scf_result;

void ref_scf_defaults(ref_scf_params* p);
/* Full SCF as HartreeFockAlgorithm::Init + Calculate (HartreeFock/HartreeFockAlgorithm.cpp:21-192).
 * D_out (n*n, or 2*n*n for UHF alpha then beta) may be NULL. */
int ref_scf(ref_mol* m, const ref_scf_params* p, ref_scf_result* r, double* D_out);

#ifdef __cplusplus
}
#endif
#endif
