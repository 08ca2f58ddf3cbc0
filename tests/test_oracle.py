"""Pin the CPU oracle before trusting it: golden vectors from an independent
McMurchie-Davidson code, the reference's own known answers, textbook values,
closed forms and invariants.  CPU only."""
import math
import os

import numpy as np
import pytest

from conftest import basis
from hartreefock_amd import geometry as g
from oracle import refcpu
import md_eri

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def load_eris(name):
    out = []
    for line in open(os.path.join(GOLD, name)):
        if line.startswith("#"):
            continue
        i, j, k, l, v = line.split()
        out.append((int(i) - 1, int(j) - 1, int(k) - 1, int(l) - 1, float(v)))
    return out


def load_matrix(name, n):
    M = np.zeros((n, n))
    for line in open(os.path.join(GOLD, name)):
        if line.startswith("#"):
            continue
        i, j, v = line.split()
        M[int(i) - 1, int(j) - 1] = M[int(j) - 1, int(i) - 1] = float(v)
    return M


# ---------------------------------------------------------------- Boys
def test_boys_matches_mpmath():
    """Oracle Boys (incomplete gamma + downward recursion, maxM = 9 as for d bases)
    against 40-digit mpmath; relative 1e-13.  The reference clamps T < 1e-10 to 1e-10
    for the top order only (HartreeFock/BoysFunction.cpp:13-14); the downward recursion
    uses the true T, so orders below the top are exact and the top order (never read
    by the vertical recursion, which needs m <= 4 Lmax) equals F_9(1e-10)."""
    rows = [l.split() for l in open(os.path.join(GOLD, "boys_mpmath.txt")) if not l.startswith("#")]
    table = {}
    for T, m, v in rows:
        table.setdefault(float(T), {})[int(m)] = float(v)
    worst = 0.0
    for T, vals in table.items():
        f = refcpu.boys(9, T)
        for m in range(10):
            ref = vals[m]
            if m == 9 and T < 1e-10:
                ref = md_eri.boys_mp(9, 1e-10)
            worst = max(worst, abs(f[m] - ref) / ref)
    assert worst < 1e-13, worst


def test_boys_closed_forms():
    f = refcpu.boys(9, 0.0)
    for m in range(9):                       # F_m(0) = 1/(2m+1) below the clamped top order
        assert abs(f[m] - 1.0 / (2 * m + 1)) < 1e-15
    for T in (0.3, 2.0, 11.0):
        f0 = refcpu.boys(9, T)[0]
        assert abs(f0 - 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))) < 1e-14


# ---------------------------------------------------------------- known answers
def test_nuclear_repulsion_known_answers():
    """HartreeFock/Test.cpp:623 and :781."""
    m = refcpu.Molecule(basis("sto3g.txt"), g.water_crawford())
    assert abs(m.nuclear_repulsion() - 8.002367061810450) < 1e-11
    m = refcpu.Molecule(basis("sto3g.txt"), g.methane_crawford())
    assert abs(m.nuclear_repulsion() - 13.497304462036480) < 1e-11


def test_crawford_water_scf():
    """Crawford Project #3 H2O/STO-3G SCF energy -74.942079928192 (the external target the
    reference's harness compares against, HartreeFock/Test.cpp:32-34)."""
    m = refcpu.Molecule(basis("sto3g.txt"), g.water_crawford())
    r, _ = m.scf(alpha=0.5, init_guess=0.0, normal_iter_after_diis=0)
    assert r.converged
    assert abs(r.energy - (-74.942079928192)) < 1e-9


def test_h2_szabo_ostlund():
    """Szabo & Ostlund STO-3G H2 at R = 1.4 bohr: S12 = 0.6593, T11 = 0.7600, T12 = 0.2365,
    E = -1.1167 Eh.  (With DIIS the reference's solver meets an exactly singular DIIS
    system for this 2-function case, so DIIS is off here.)"""
    m = refcpu.Molecule(basis("sto3g.txt"), [(1, (0, 0, 0)), (1, (0, 0, 1.4))])
    S, T, V = m.one_electron()
    assert abs(S[0, 1] - 0.6593) < 1e-4 and abs(T[0, 0] - 0.7600) < 1e-4 and abs(T[0, 1] - 0.2365) < 1e-4
    assert abs(V[0, 0] - (-1.8804)) < 1e-4 and abs(V[0, 1] - (-1.1948)) < 1e-4
    assert abs(m.eri(0, 0, 0, 0) - 0.7746) < 1e-4 and abs(m.eri(0, 0, 1, 1) - 0.5697) < 1e-4
    assert abs(m.eri(1, 0, 0, 0) - 0.4441) < 1e-4 and abs(m.eri(1, 0, 1, 0) - 0.2970) < 1e-4
    r, _ = m.scf(alpha=0.5, init_guess=0.0, normal_iter_after_diis=0, use_diis=0)
    assert abs(r.energy - (-1.1167)) < 1e-4


def test_h_atom_uhf():
    """H atom, STO-3G, UHF: E = <1s|h|1s> = -0.4666 Eh (Szabo & Ostlund), <S^2> = 3/4."""
    m = refcpu.Molecule(basis("sto3g.txt"), [(1, (0, 0, 0))])
    r, _ = m.scf(unrestricted=True, normal_iter_after_diis=0)
    assert abs(r.energy - (-0.4666)) < 1e-4
    assert abs(r.s2 - 0.75) < 1e-12


# ---------------------------------------------------------------- golden vectors
def test_eri_h2o_sto3g_all_unique_vs_md():
    m = refcpu.Molecule(basis("sto3g.txt"), g.water_cfg1())
    packed = m.eri_packed()
    gold = load_eris("eri_h2o_sto3g.txt")
    assert len(gold) == 406
    err = max(abs(packed[refcpu.packed_index(i, j, k, l)] - v) for i, j, k, l, v in gold)
    assert err < 1e-13, err


def test_eri_h2o_631gs_sample_vs_md():
    m = refcpu.Molecule(basis("6-31g_st_.1.nw"), g.water_cfg1())
    for i, j, k, l, v in load_eris("eri_h2o_631gs_sample.txt"):
        assert abs(m.eri(i, j, k, l) - v) < 1e-13


@pytest.mark.parametrize("bname,tag", [("sto3g.txt", "sto3g"), ("6-31g_st_.1.nw", "631gs")])
def test_one_electron_vs_md(bname, tag):
    m = refcpu.Molecule(basis(bname), g.water_cfg1())
    S, T, V = m.one_electron()
    n = m.nbf
    for name, M in (("s", S), ("t", T), ("v", V)):
        G = load_matrix("%s_h2o_%s.txt" % (name, tag), n)
        assert np.abs(M - G).max() < 1e-12, name


def test_dddd_transfer_convention():
    """The reference fills electron-transfer rows |e| >= |f|-1 only
    (HartreeFock/GaussianTwoElectrons.cpp:152).  (dd|dd) needs |e|=2 at |f|=4, whose
    coefficient in the bra HRR is proportional to AB^2: with a one-centre bra the
    reference convention is exact, with a two-centre bra it is not.  Exact mode
    reproduces the independent MD values everywhere."""
    atoms = [(6, (0.0, 0.0, 0.0)), (6, (0.3, 1.2, 2.5))]
    m = refcpu.Molecule(basis("6-31g_st_.1.nw"), atoms)
    _, _, at, _ = m.bf_info()
    gold = load_eris("eri_c2_631gs_dddd.txt")
    two_centre_diff = 0.0
    for i, j, k, l, v in gold:
        refcpu.set_exact_transfer(True)
        ex = m.eri(i, j, k, l)
        refcpu.set_exact_transfer(False)
        rv = m.eri(i, j, k, l)
        assert abs(ex - v) < 1e-13
        if at[i] == at[j] and at[k] == at[l]:
            assert abs(rv - v) < 1e-13             # both pairs one-centre: no omitted term survives
        else:
            two_centre_diff = max(two_centre_diff, abs(rv - v))
    assert two_centre_diff > 1e-4


# ---------------------------------------------------------------- invariants
def test_eri_permutational_symmetry_and_schwarz():
    m = refcpu.Molecule(basis("6-31g_st_.1.nw"), g.water_cfg1())
    rng = np.random.default_rng(3)
    n = m.nbf
    for _ in range(40):
        i, j, k, l = (int(x) for x in rng.integers(0, n, size=4))
        v = m.eri(i, j, k, l)
        for p in ((j, i, k, l), (i, j, l, k), (k, l, i, j), (l, k, j, i)):
            assert abs(m.eri(*p) - v) < 1e-14
        assert abs(v) <= math.sqrt(m.eri(i, j, i, j) * m.eri(k, l, k, l)) + 1e-14
        assert m.eri(i, j, i, j) >= -1e-15


def test_eri_translation_invariance():
    atoms = g.water_cfg1()
    shifted = [(z, (x + 0.37, y - 1.1, w + 2.0)) for z, (x, y, w) in atoms]
    a = refcpu.Molecule(basis("6-31g_st_.1.nw"), atoms).eri_packed()
    b = refcpu.Molecule(basis("6-31g_st_.1.nw"), shifted).eri_packed()
    assert np.abs(a - b).max() < 1e-12


def test_jk_from_packed_matches_dense_contraction():
    m = refcpu.Molecule(basis("sto3g.txt"), g.water_cfg1())
    n = m.nbf
    packed = m.eri_packed()
    full = np.zeros((n, n, n, n))
    for i in range(n):
        for j in range(n):
            for k in range(n):
                for l in range(n):
                    full[i, j, k, l] = packed[refcpu.packed_index(i, j, k, l)]
    rng = np.random.default_rng(1)
    X = rng.normal(size=(n, n))
    D = X + X.T
    J, K = refcpu.jk_from_packed(n, packed, D)
    assert np.abs(J - np.einsum("ijkl,kl->ij", full, D)).max() < 1e-12
    assert np.abs(K - np.einsum("ilkj,kl->ij", full, D)).max() < 1e-12


def test_scf_energy_is_variational_minimum_and_trace():
    """RHF density: Tr(D S) = N_e; converged energy below the core-guess energy."""
    m = refcpu.Molecule(basis("sto3g.txt"), g.water_cfg1())
    r, D = m.scf(alpha=0.5, init_guess=0.0, normal_iter_after_diis=0)
    S, T, V = m.one_electron()
    assert abs(np.trace(D[0] @ S) - 10.0) < 1e-8
    assert r.converged and r.energy < -74.9


# ---------------------------------------------------------------- workload sizes
def test_config_basis_sizes():
    """Basis sizes of the benchmark workloads (n = 7 / 194 / 800 / 462)."""
    assert refcpu.Molecule(basis("sto3g.txt"), g.water_cfg1()).nbf == 7
    assert refcpu.Molecule(basis("6-31g_st_.1.nw"), g.decane_cfg2()).nbf == 194
    assert refcpu.Molecule(basis("6-31g_st__st_.nw"), g.water32_cfg3()).nbf == 800
    assert refcpu.Molecule(basis("6-31+g_st_.1.nw"), g.eicosyl_cfg5()).nbf == 462
    e = refcpu.Molecule(basis("6-31+g_st_.1.nw"), g.eicosyl_cfg5()).electrons()
    assert e == (80, 81)


def test_serving_mix_sizes():
    mols = g.serving_mix(60, seed=1)
    sizes = [refcpu.Molecule(basis("6-31g.1.nw"), atoms).nbf for _, atoms in mols]
    assert min(sizes) >= 13 and max(sizes) <= 212


def test_geometry_generators():
    atoms = g.alkane_zigzag(10, seed=0, jitter=0.0)
    c = np.array([p for z, p in atoms if z == 6]) * g.BOHR_ANGSTROM
    d = np.linalg.norm(np.diff(c, axis=0), axis=1)
    assert np.allclose(d, 1.54, atol=1e-12)
    w = g.water_cluster(32, seed=2024, box=10.0)
    o = np.array([p for z, p in w if z == 8]) * g.BOHR_ANGSTROM
    dist = np.linalg.norm(o[:, None] - o[None], axis=-1) + np.eye(len(o)) * 99
    assert dist.min() >= 2.6 - 1e-9


# ---------------------------------------------------------------- properties and MP2
def _crawford_water_state():
    m = refcpu.Molecule(basis("sto3g.txt"), g.water_crawford())
    r, D = m.scf(alpha=0.5, init_guess=0.0, normal_iter_after_diis=0)
    assert r.converged
    return m, D[0]


def test_crawford_water_properties_known_answers():
    """Crawford Project #3, H2O/STO-3G: Mulliken charges O -0.253146052405, H +0.126573026202 and the
    dipole moment 0.603521296525 au along y (the reference quotes 0.603521 au for the same molecule,
    HartreeFock/Test.cpp:660-661)."""
    m, D = _crawford_water_state()
    S = m.one_electron()[0]
    L, lmn, atom_of_bf, sh = m.bf_info()
    atoms = g.water_crawford()
    Z = [a[0] for a in atoms]
    R = [a[1] for a in atoms]
    q = refcpu.mulliken_charges(D, S, Z, atom_of_bf)
    assert np.abs(q - np.array([-0.253146052405, 0.126573026202, 0.126573026202])).max() < 1e-8
    mu = refcpu.dipole(D, m.moment(), Z, R)
    assert abs(mu[0]) < 1e-10 and abs(mu[2]) < 1e-10
    assert abs(mu[1] - 0.603521296525) < 1e-8


def test_moment_integrals_translation():
    """<a|r|b> about the origin shifts by t S_ab when the molecule moves by t; the dipole of a neutral
    molecule is unchanged."""
    atoms = g.water_cluster(1, seed=3)
    t = np.array([0.7, -1.3, 2.1])
    moved = [(z, tuple(np.add(xyz, t))) for z, xyz in atoms]
    m0 = refcpu.Molecule(basis("6-31g_st_.1.nw"), atoms)
    m1 = refcpu.Molecule(basis("6-31g_st_.1.nw"), moved)
    S = m0.one_electron()[0]
    for a, b, tk in zip(m0.moment(), m1.moment(), t):
        assert np.abs(b - (a + tk * S)).max() < 1e-12


def test_crawford_water_mp2_known_answer():
    """Crawford Project #4, H2O/STO-3G: MP2 correlation energy -0.049149636120 Eh (all electrons
    correlated, as RestrictedHartreeFock::CalculateMp2Energy, HartreeFock/RestrictedHartreeFock.cpp:251-288)."""
    m, D = _crawford_water_state()
    n = m.nbf
    S, T, V = m.one_electron()
    e2 = refcpu.mp2_rhf(n, m.eri_packed(), D, T + V, S, m.electrons()[0])
    assert abs(e2 - (-0.049149636120)) < 1e-9


def test_uhf_mp2_reference_formula_closed_shell_limit():
    """The reference's UHF MP2 formula (same-spin (ia|jb)^2 sums, times 1/4) on a closed shell with
    equal alpha and beta densities equals half the RHF sum of (ia|jb)^2 / denominators (its formula
    has no exchange and no opposite-spin term).  Pins the spin bookkeeping of the oracle's UHF MP2
    against the RHF orbitals; the UHF formula itself is the reference's as written."""
    m, D = _crawford_water_state()
    n = m.nbf
    S, T, V = m.one_electron()
    eri = m.eri_packed()
    nocc = m.electrons()[0]
    e_u = refcpu.mp2_uhf_reference(n, eri, 0.5 * D, 0.5 * D, T + V, S, nocc, nocc)
    J, K = refcpu.jk_from_packed(n, eri, D)
    import scipy.linalg
    eps, C = scipy.linalg.eigh(T + V + J - 0.5 * K, S)
    G = refcpu.unpack_eri(n, eri)
    ov = np.einsum("pqrs,pi,qa,rj,sb->iajb", G, C[:, :nocc], C[:, nocc:], C[:, :nocc], C[:, nocc:])
    eo, ev = eps[:nocc], eps[nocc:]
    den = eo[:, None, None, None] - ev[None, :, None, None] + eo[None, None, :, None] - ev[None, None, None, :]
    assert abs(e_u - 0.5 * np.sum(ov * ov / den)) < 1e-13


# ---------------------------------------------------------------- UHF, (dd|dd) convention, field: closed relations
@pytest.mark.parametrize("bname,atoms", [("sto3g.txt", g.water_cfg1()), ("6-31g_st_.1.nw", g.alkane_zigzag(2, seed=5))],
                         ids=["h2o_sto3g", "ethane_631gs"])
def test_closed_shell_uhf_equals_rhf(bname, atoms):
    """A closed shell in UHF without the iteration-0 asymmetry has D_alpha = D_beta = D/2, so
    F_alpha = h + J[D_alpha + D_beta] - K[D_alpha] is the RHF Fock matrix: the energy must equal RHF's and
    <S^2> must vanish.  This pins the alpha/beta Coulomb coupling of the UHF Fock build
    (HartreeFock/UnrestrictedHartreeFock.cpp:228-238; dropping D_beta J changes the energy by Eh) and the
    overlap term of <S^2> (Tr(P_a S P_b S) = N_alpha here)."""
    kw = dict(alpha=0.5, init_guess=0.0, normal_iter_after_diis=0)
    r, _ = refcpu.Molecule(basis(bname), atoms).scf(**kw)
    u, _ = refcpu.Molecule(basis(bname), atoms).scf(unrestricted=True, add_asymmetry=0, **kw)
    assert r.converged and u.converged
    assert abs(r.energy - u.energy) <= 1e-10, (r.energy, u.energy)
    assert abs(u.s2) <= 1e-10


def test_dddd_reference_convention_error_scales_as_ab_squared():
    """The reference's transfer range leaves the (dd|dd) intermediates |e| = 2 at |f| = 4 at zero
    (HartreeFock/GaussianTwoElectrons.cpp:152); their weight in the bra HRR is the square of the bra's centre
    separation, so ref - exact for a two-centre bra must fall by 4 each time the separation halves (a range
    starting one level off would leave a term linear in AB: ratio 2)."""
    d = np.array([0.3, 1.2, 2.5])
    d /= np.linalg.norm(d)
    diffs = []
    for R in (0.04, 0.02, 0.01):
        m = refcpu.Molecule(basis("6-31g_st_.1.nw"), [(6, (0.0, 0.0, 0.0)), (6, tuple(R * d))])
        L, _, atom, _ = m.bf_info()
        dA = [i for i in range(m.nbf) if L[i] == 2 and atom[i] == 0]
        dB = [i for i in range(m.nbf) if L[i] == 2 and atom[i] == 1]
        worst = 0.0
        try:
            for i in dA:
                for j in dB:
                    for k in dA[:3]:
                        for l in dA[:3]:
                            refcpu.set_exact_transfer(True)
                            ex = m.eri(i, j, k, l)
                            refcpu.set_exact_transfer(False)
                            worst = max(worst, abs(m.eri(i, j, k, l) - ex))
        finally:
            refcpu.set_exact_transfer(False)
        diffs.append(worst)
    assert diffs[2] > 0
    for a, b in zip(diffs, diffs[1:]):
        assert abs(a / b - 4.0) < 0.01, diffs


def test_field_finite_difference_dipole_oracle():
    """The reference's dipole check (HartreeFock/Test.cpp:626-733) on the oracle: H2O/STO-3G (Crawford geometry,
    no DIIS, alpha 0.5, initGuess 0.7) in fields of +-0.001 au, h = T + V + E . M
    (HartreeFock/HartreeFockAlgorithm.cpp:38).  The nuclear moment plus the central difference of the energy
    equals the analytic SCF dipole within 2e-6 au, both 0.60352 au (the reference's comment quotes 0.603521
    and 0.60352); a wrong sign of the field term flips the difference."""
    atoms = g.water_crawford()
    kw = dict(use_diis=0, alpha=0.5, init_guess=0.7, normal_iter_after_diis=0)
    mol = refcpu.Molecule(basis("sto3g.txt"), atoms)
    r, D = mol.scf(**kw)
    mu = refcpu.dipole(D[0], mol.moment(), [a[0] for a in atoms], [a[1] for a in atoms])
    d = 0.001
    fd = np.array([z * np.asarray(x, float) for z, x in atoms]).sum(axis=0)
    for k in range(3):
        E = np.zeros(3)
        E[k] = -d
        e1 = refcpu.Molecule(basis("sto3g.txt"), atoms).set_field(E).scf(**kw)[0].energy
        E[k] = d
        e2 = refcpu.Molecule(basis("sto3g.txt"), atoms).set_field(E).scf(**kw)[0].energy
        fd[k] += (e1 - e2) / (2 * d)
    assert np.abs(mu - fd).max() <= 2e-6, (mu, fd)
    assert abs(np.linalg.norm(mu) - 0.603521) <= 1e-6
    assert abs(np.linalg.norm(fd) - 0.60352) <= 1e-5


def test_ion_and_high_spin_electron_counts():
    """Molecule::alphaElectrons / betaElectrons (HartreeFock/Molecule.h:16-17): O2 in STO-3G as a triplet
    (9 alpha, 7 beta) has <S^2> just above S(S+1) = 2 and lies below the closed-shell RHF singlet; OH- in RHF
    has 5 doubly occupied orbitals, so its energy differs from neutral OH's."""
    kw = dict(alpha=0.5, init_guess=0.0, normal_iter_after_diis=0)
    o2 = [(8, (0.0, 0.0, 0.0)), (8, (0.0, 0.0, 2.28))]
    t, _ = refcpu.Molecule(basis("sto3g.txt"), o2).set_electrons(9, 7).scf(unrestricted=True, **kw)
    s, _ = refcpu.Molecule(basis("sto3g.txt"), o2).scf(**kw)
    assert t.converged and s.converged
    assert 2.0 <= t.s2 < 2.1 and t.energy < s.energy
    oh = [(8, (0.0, 0.0, 0.0)), (1, (0.0, 0.0, 1.83))]
    m = refcpu.Molecule(basis("sto3g.txt"), oh).set_electrons(5, 5)
    assert m.electrons() == (5, 5)
    a, _ = m.scf(**kw)
    assert a.converged
