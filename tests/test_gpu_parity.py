"""GPU parity against the CPU oracle: ERIs (1e-12 abs), J/K, one-electron, SCF energies (1e-8 Eh)."""
import numpy as np
import pytest

from conftest import basis
from hartreefock_amd import geometry as g
from oracle import refcpu

pytestmark = pytest.mark.gpu

ERI_TOL = 1e-12      # absolute, ERIs
E_TOL = 1e-8         # Eh, SCF energies

SMALL = [
    ("h2o_sto3g_cfg1", "sto3g.txt", g.water_cfg1()),
    ("h2o_crawford_sto3g", "sto3g.txt", g.water_crawford()),
    ("ch4_crawford_sto3g", "sto3g.txt", g.methane_crawford()),
    ("h2o_631gs", "6-31g_st_.1.nw", g.water_cfg1()),
    ("ch4_631gs", "6-31g_st_.1.nw", g.methane_crawford()),
    ("ethane_631gs", "6-31g_st_.1.nw", g.alkane_zigzag(2, seed=5)),
    ("h2o_631gss", "6-31g_st__st_.nw", g.water_cluster(1, seed=3)),
]


def _print_case(name):
    print("case", name, flush=True)


@pytest.mark.parametrize("name,bfile,atoms", SMALL, ids=[s[0] for s in SMALL])
@pytest.mark.parametrize("exact", [0, 1])
def test_packed_eri_matches_oracle(gpu_engine_factory, name, bfile, atoms, exact):
    _print_case(name)
    eng = gpu_engine_factory(exact_transfer=exact).set_molecule(basis(bfile), atoms)
    mol = refcpu.Molecule(basis(bfile), atoms)
    assert eng.nbf == mol.nbf
    refcpu.set_exact_transfer(bool(exact))
    try:
        ref = mol.eri_packed()
    finally:
        refcpu.set_exact_transfer(False)
    got = eng.eri_packed()
    n = eng.nbf
    m = refcpu.packed_index(n - 1, n - 1, n - 1, n - 1) + 1
    err = np.abs(got[:m] - ref[:m]).max()
    assert err <= ERI_TOL, (name, err)


@pytest.mark.parametrize("name,bfile,atoms", SMALL[:6], ids=[s[0] for s in SMALL[:6]])
def test_one_electron_matches_oracle(gpu_engine_factory, name, bfile, atoms):
    _print_case(name)
    eng = gpu_engine_factory().set_molecule(basis(bfile), atoms)
    mol = refcpu.Molecule(basis(bfile), atoms)
    for a, b in zip(eng.one_electron(), mol.one_electron()):
        assert np.abs(a - b).max() <= 1e-12
    assert abs(eng.nuclear_repulsion() - mol.nuclear_repulsion()) <= 1e-12


@pytest.mark.parametrize("name,bfile,atoms", SMALL[3:6], ids=[s[0] for s in SMALL[3:6]])
@pytest.mark.parametrize("ndens", [1, 2])
def test_jk_matches_oracle(gpu_engine_factory, name, bfile, atoms, ndens):
    _print_case(name)
    eng = gpu_engine_factory(schwarz_threshold=0.0).set_molecule(basis(bfile), atoms)
    mol = refcpu.Molecule(basis(bfile), atoms)
    n = eng.nbf
    eri = mol.eri_packed()
    rng = np.random.default_rng(7)
    Ds = []
    for _ in range(ndens):
        X = rng.normal(size=(n, n))
        Ds.append(0.5 * (X + X.T))
    D = np.array(Ds) if ndens == 2 else Ds[0]
    J, K = eng.build_jk(D)
    Jref = sum(refcpu.jk_from_packed(n, eri, d)[0] for d in Ds)
    assert np.abs(J - Jref).max() <= 1e-10 * max(1.0, np.abs(Jref).max())
    Kg = K if ndens == 2 else [K]
    for d, k in zip(Ds, Kg):
        Kref = refcpu.jk_from_packed(n, eri, d)[1]
        assert np.abs(k - Kref).max() <= 1e-10 * max(1.0, np.abs(Kref).max())


RHF_CASES = [
    ("h2o_sto3g_cfg1", "sto3g.txt", g.water_cfg1(), dict(alpha=0.5, init_guess=0.0)),
    ("h2o_crawford_sto3g", "sto3g.txt", g.water_crawford(), dict(alpha=0.5, init_guess=0.7)),
    ("ch4_crawford_sto3g", "sto3g.txt", g.methane_crawford(), dict()),
    ("h2o_631gs", "6-31g_st_.1.nw", g.water_cfg1(), dict(alpha=0.5, init_guess=0.0)),
    ("ethane_631gs", "6-31g_st_.1.nw", g.alkane_zigzag(2, seed=5), dict(alpha=0.5, init_guess=0.0)),
]


@pytest.mark.parametrize("name,bfile,atoms,kw", RHF_CASES, ids=[c[0] for c in RHF_CASES])
def test_rhf_energy_matches_oracle(gpu_engine_factory, name, bfile, atoms, kw):
    _print_case(name)
    kw = dict(kw, normal_iter_after_diis=0)
    eng = gpu_engine_factory(schwarz_threshold=0.0).set_molecule(basis(bfile), atoms)
    r = eng.scf(**kw)
    ref, _ = refcpu.Molecule(basis(bfile), atoms).scf(**kw)
    assert r.converged and ref.converged
    assert abs(r.energy - ref.energy) <= E_TOL, (r.energy, ref.energy)


UHF_CASES = [
    ("oh_sto3g", "sto3g.txt", [(8, (0.0, 0.0, 0.0)), (1, (0.0, 0.0, 1.83))]),
    ("ch3_631gs", "6-31g_st_.1.nw", [(6, (0.0, 0.0, 0.0)), (1, (2.04, 0.0, 0.0)), (1, (-1.02, 1.767, 0.0)),
                                     (1, (-1.02, -1.767, 0.0))]),
]


@pytest.mark.parametrize("name,bfile,atoms", UHF_CASES, ids=[c[0] for c in UHF_CASES])
def test_uhf_energy_matches_oracle(gpu_engine_factory, name, bfile, atoms):
    _print_case(name)
    kw = dict(alpha=0.5, init_guess=0.0, normal_iter_after_diis=0)
    eng = gpu_engine_factory(schwarz_threshold=0.0).set_molecule(basis(bfile), atoms)
    r = eng.scf(unrestricted=True, **kw)
    ref, _ = refcpu.Molecule(basis(bfile), atoms).scf(unrestricted=True, **kw)
    assert r.converged and ref.converged
    assert abs(r.energy - ref.energy) <= E_TOL, (r.energy, ref.energy)
    assert abs(r.s2 - ref.s2) <= 1e-6
    assert 0.75 <= r.s2 < 0.80          # doublet with small spin contamination
