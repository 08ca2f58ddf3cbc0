"""GPU parity and SCF invariants at the benchmark configurations' full sizes:
C10H22/6-31G* RHF (194 functions), (H2O)32/6-31G** RHF (800) and the C20H41/6-31+G* UHF doublet (462, diffuse sp).

ERIs: the direct J/K build of a density with one symmetric non-zero pair (k, l) returns ERIs, J_ij = sum_mn
(ij|mn) D_mn = (2 - delta_kl) (ij|kl) and K_ij = (ik|lj) + (il|kj) (one term when k = l), so each build at the
molecule's full size yields n^2 integrals of every bra class against the ket (k, l).  Six (k, l) pair classes times
six bra pair classes sampled cover all 21 quartet classes; ~13 000 ERIs per configuration and convention are
compared with the oracle's IntegralsRepository::getElectronElectron restatement (HartreeFock/IntegralsRepository.cpp:
372-433) at 1e-12 absolute, in the reference's (dd|dd) transfer convention and in the exact one.

SCF: each configuration's SCF on the GPU must satisfy what its converged state fixes whatever the route:
Tr(D S) = N per spin, D S D = D (per spin), the commutator F D S - S D F ~ 0 with F rebuilt from the GPU's J/K of D,
and the energy functional E[D] equal to the reported energy; direct and in-core J/K of the converged density agree.
"""
import numpy as np
import pytest

from conftest import basis
from hartreefock_amd import geometry as g
from oracle import refcpu

pytestmark = pytest.mark.gpu

CFGS = [
    ("cfg2_decane_631gs", "6-31g_st_.1.nw", g.decane_cfg2, 194, False),
    ("cfg3_water32_631gss", "6-31g_st__st_.nw", g.water32_cfg3, 800, False),
    ("cfg5_eicosyl_631pgs", "6-31+g_st_.1.nw", g.eicosyl_cfg5, 462, True),
]
SCF_KW = dict(alpha=0.5, init_guess=0.0, normal_iter_after_diis=0, energy_tolerance=1e-8, density_tolerance=1e-6,
              diis_tolerance=1e-6)
PAIR_CLASSES = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]


def _print_case(*a):
    print("case", *a, flush=True)


def _pick(rng, idx, n):
    return idx[rng.integers(0, len(idx), size=n)]


@pytest.mark.parametrize("route", ["direct", "incore"])
@pytest.mark.parametrize("exact", [0, 1], ids=["reference_dddd", "exact"])
@pytest.mark.parametrize("name,bfile,make,nbf,uhf", CFGS, ids=[c[0] for c in CFGS])
def test_sampled_eris_through_jk(gpu_engine_factory, name, bfile, make, nbf, uhf, exact, route):
    """Screening at the configurations' Schwarz threshold (1e-12), so the task lists are the production ones; a
    screened quartet's integrals are below the threshold, within the 1e-12 tolerance.  The in-core route
    (n <= 512) evaluates the pair-matrix store and contracts it."""
    if route == "incore" and nbf > 512:
        pytest.skip("the in-core store takes n <= 512")
    _print_case(name, "exact" if exact else "reference convention", route)
    atoms = make()
    eng = gpu_engine_factory(schwarz_threshold=1e-12, exact_transfer=exact).set_molecule(basis(bfile), atoms)
    build = eng.build_jk_incore if route == "incore" else eng.build_jk
    mol = refcpu.Molecule(basis(bfile), atoms)
    n = eng.nbf
    assert n == nbf == mol.nbf
    L, _, atom, _ = mol.bf_info()
    byL = {l: np.where(L == l)[0] for l in range(3)}
    rng = np.random.default_rng(1000 + 10 * nbf + exact)
    refcpu.set_exact_transfer(bool(exact))
    worst, count, classes, two_centre_dddd = 0.0, 0, set(), 0
    try:
        for lk, ll in PAIR_CLASSES:
            k = int(_pick(rng, byL[lk], 1)[0])
            # half the kets one-centre, half two-centre (a d ket on the same atom as a d bra function is
            # where the reference's (dd|dd) convention departs from the exact integrals)
            same = [x for x in byL[ll] if atom[x] == atom[k] and x != k] if len(classes) % 2 == 0 else []
            l = int(same[0]) if same else int(_pick(rng, byL[ll], 1)[0])
            D = np.zeros((n, n))
            D[k, l] = D[l, k] = 1.0
            J, K = build(D)
            fac = 1.0 if k == l else 2.0
            for li, lj in PAIR_CLASSES:
                ii = _pick(rng, byL[li], 150)
                jj = _pick(rng, byL[lj], 150)
                for i, j in zip(ii.tolist(), jj.tolist()):
                    ref = mol.eri(i, j, k, l)
                    worst = max(worst, abs(J[i, j] / fac - ref))
                    kref = mol.eri(i, k, l, j) + (mol.eri(i, l, k, j) if k != l else 0.0)
                    worst = max(worst, abs(K[i, j] - kref) / (2.0 if k != l else 1.0))
                    count += 3 if k != l else 2
                    classes.add(tuple(sorted([(li, lj), (lk, ll)])))
                    if (li, lj, lk, ll) == (2, 2, 2, 2) and atom[i] != atom[j]:
                        two_centre_dddd += 1
    finally:
        refcpu.set_exact_transfer(False)
    print("%s: %d ERIs compared, %d quartet classes, %d two-centre-bra (dd|dd), worst |d| %.2e"
          % (name, count, len(classes), two_centre_dddd, worst), flush=True)
    assert len(classes) == 21
    assert count >= 10000 and two_centre_dddd > 0
    assert worst <= 1e-12, worst


def _fock(J, K, h, uhf, Da=None):
    return [h + J - K[0], h + J - K[1]] if uhf else [h + J - 0.5 * K]


@pytest.mark.parametrize("name,bfile,make,nbf,uhf", CFGS, ids=[c[0] for c in CFGS])
def test_config_scf_invariants(gpu_engine_factory, name, bfile, make, nbf, uhf):
    _print_case(name)
    atoms = make()
    eng = gpu_engine_factory(schwarz_threshold=1e-12, exact_transfer=1).set_molecule(basis(bfile), atoms)
    n = eng.nbf
    incore = n <= 512
    r, D = eng.scf(unrestricted=uhf, want_density=True, incore=1 if incore else 0,
                   max_iterations=400 if uhf else 300, **SCF_KW)
    print("%s: E %.10f iterations %d converged %d incore %d s2 %.6f" % (name, r.energy, r.iterations, r.converged,
                                                                      r.incore, r.s2), flush=True)
    assert r.converged and r.incore == (1 if incore else 0)
    S, T, V = eng.one_electron()
    h = T + V
    na, nb = eng.electrons()
    Ds = [D[0], D[1]] if uhf else [D[0]]
    occ = [na, nb] if uhf else [na + nb]
    # electron count and idempotency per spin (RHF: D = 2 P with P S P = P)
    for Dk, nk in zip(Ds, occ):
        assert abs(np.sum(Dk * S.T) - nk) <= 1e-8, (np.sum(Dk * S.T), nk)
        scale = 1.0 if uhf else 2.0
        assert np.abs(Dk @ S @ Dk - scale * Dk).max() <= 1e-6
    # Fock matrices from the GPU's direct J/K of the converged density: the DIIS commutator vanishes and the
    # energy functional reproduces the SCF energy.  With the in-core store, the same from its J/K, and the two
    # builds against each other.  Every figure is printed before any is asserted, so a failure shows which build
    # disagrees.
    Enuc = eng.nuclear_repulsion()
    Dx = np.array(Ds) if uhf else Ds[0]

    def fock_checks(J, K):
        Fs = _fock(J, K, h, uhf)
        comm = max(np.abs(Fk @ Dk @ S - S @ Dk @ Fk).max() for Fk, Dk in zip(Fs, Ds))
        return comm, sum(0.5 * np.sum(Dk * (h + Fk)) for Dk, Fk in zip(Ds, Fs)) + Enuc

    J, K = eng.build_jk(Dx)
    J, K = np.asarray(J), np.asarray(K)
    comm, E = fock_checks(J, K)
    print("%s: direct J/K: max|F D S - S D F| %.2e, E[D] %.10f, reported %.10f" % (name, comm, E, r.energy),
          flush=True)
    if incore:
        # The two sum the same ERIs (same screened quartets) in different orders, so they differ by rounding, which
        # scales with the terms' magnitudes rather than with their sums: the diffuse 6-31+G* density has large
        # elements of both signs.  The scale is the J/K of |D| (the same contraction without cancellation).
        Ji, Ki = eng.build_jk_incore(Dx)
        Ji, Ki = np.asarray(Ji), np.asarray(Ki)
        comm_i, E_i = fock_checks(Ji, Ki)
        Ja, Ka = eng.build_jk(np.abs(Dx))
        dj = np.abs(Ji - J).max() / np.abs(Ja).max()
        dk = np.abs(Ki - K).max() / np.abs(np.asarray(Ka)).max()
        rel = max(np.abs(Ji - J).max() / np.abs(J).max(), np.abs(Ki - K).max() / np.abs(K).max())
        print("%s: in-core J/K: max|F D S - S D F| %.2e, E[D] %.10f; direct vs in-core: max|dJ| / max J[|D|] %.2e, "
              "max|dK| / max K[|D|] %.2e (relative to J, K themselves %.2e; max|D| %.3f)"
              % (name, comm_i, E_i, dj, dk, rel, max(np.abs(x).max() for x in Ds)), flush=True)
    assert comm <= 1e-5
    assert abs(E - r.energy) <= 1e-6
    if uhf:
        assert 0.75 <= r.s2 < 0.80
    if incore:
        assert dj <= 1e-12 and dk <= 1e-12 and rel <= 1e-9
    if name.startswith("cfg2"):
        # the direct SCF reaches the in-core one's energy
        rd = eng.scf(incore=0, **SCF_KW)
        assert rd.converged and abs(rd.energy - r.energy) <= 1e-7, (rd.energy, r.energy)
