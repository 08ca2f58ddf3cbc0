"""Deterministic mode (hf_options.deterministic = 1): the J/K builds (direct and in-core, FP64 and the mixed tier's
FP32 quartets) and the nuclear-attraction matrix accumulate in a fixed-point form whose integer sums do not depend
on the order of the GPU's additions.  Two builds, two contexts and two SCF runs must give the same bits; the FP64
atomic mode's results must agree with them to rounding (J/K 1e-12 relative, energies 1e-9 Eh)."""
import numpy as np
import pytest

from conftest import basis
from hartreefock_amd import geometry as g

pytestmark = pytest.mark.gpu

CASES = [
    ("pentane_631gs", "6-31g_st_.1.nw", g.alkane_zigzag(5, seed=31)),      # 99 functions, d shells
    ("water6_631gss", "6-31g_st__st_.nw", g.water_cluster(6, seed=32)),    # 150 functions
]


def _print_case(*a):
    print("case", *a, flush=True)


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(1.0, np.abs(np.asarray(b)).max())


@pytest.mark.parametrize("name,bfile,atoms", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("route,precision", [("direct", 0), ("direct", 1), ("incore", 0)],
                         ids=["direct_fp64", "direct_mixed", "incore"])
def test_deterministic_jk_bit_identical(gpu_engine_factory, name, bfile, atoms, route, precision):
    _print_case(name, route, precision)
    kw = dict(schwarz_threshold=1e-12, precision=precision)
    det = gpu_engine_factory(deterministic=1, **kw).set_molecule(basis(bfile), atoms)
    det2 = gpu_engine_factory(deterministic=1, **kw).set_molecule(basis(bfile), atoms)
    ref = gpu_engine_factory(**kw).set_molecule(basis(bfile), atoms)
    n = det.nbf
    rng = np.random.default_rng(5 + n)
    X = rng.normal(size=(2, n, n)) * 0.1
    D2 = X + X.transpose(0, 2, 1)
    fn = "build_jk_incore" if route == "incore" else "build_jk"
    for D in (D2[0], D2):                            # one density (RHF), two (UHF)
        J1, K1 = getattr(det, fn)(D)
        J2, K2 = getattr(det, fn)(D)
        J3, K3 = getattr(det2, fn)(D)
        assert np.array_equal(J1, J2) and np.array_equal(K1, K2)
        assert np.array_equal(J1, J3) and np.array_equal(K1, K3)
        Jr, Kr = getattr(ref, fn)(D)
        print("%s %s nd %d: det vs atomic rel |dJ| %.2e |dK| %.2e" % (name, route, D.ndim - 1, _rel(J1, Jr),
                                                                     _rel(K1, Kr)), flush=True)
        assert _rel(J1, Jr) <= 1e-12 and _rel(K1, Kr) <= 1e-12
    V1 = det.one_electron()[2]
    V2 = det2.one_electron()[2]
    assert np.array_equal(V1, V2)
    assert _rel(V1, ref.one_electron()[2]) <= 1e-13


SCF_KW = dict(alpha=0.5, init_guess=0.0, normal_iter_after_diis=0, energy_tolerance=1e-10, density_tolerance=1e-8,
              diis_tolerance=1e-8, max_iterations=200)


@pytest.mark.parametrize("incore", [0, 1], ids=["direct", "incore"])
@pytest.mark.parametrize("name,bfile,atoms,uhf", [
    ("pentane_631gs", "6-31g_st_.1.nw", g.alkane_zigzag(5, seed=31), False),
    ("pentyl_631gs_uhf", "6-31g_st_.1.nw", g.alkane_zigzag(5, seed=33, drop_last_h=True), True),
], ids=["pentane_rhf", "pentyl_uhf"])
def test_deterministic_scf_bit_identical(gpu_engine_factory, name, bfile, atoms, uhf, incore):
    _print_case(name, incore)
    runs = []
    for _ in range(2):
        eng = gpu_engine_factory(schwarz_threshold=1e-12, deterministic=1).set_molecule(basis(bfile), atoms)
        r, D = eng.scf(unrestricted=uhf, want_density=True, incore=incore, **SCF_KW)
        assert r.converged
        runs.append((r.energy, r.iterations, D))
        eng.close()
    assert runs[0][0] == runs[1][0] and runs[0][1] == runs[1][1]
    assert np.array_equal(runs[0][2], runs[1][2])
    ref = gpu_engine_factory(schwarz_threshold=1e-12).set_molecule(basis(bfile), atoms)
    rr = ref.scf(unrestricted=uhf, incore=incore, **SCF_KW)
    print("%s: deterministic E %.12f (%d iterations), atomic E %.12f" % (name, runs[0][0], runs[0][1], rr.energy),
          flush=True)
    assert rr.converged and abs(rr.energy - runs[0][0]) <= 1e-9


def test_deterministic_scf_bit_identical_water32(gpu_engine_factory):
    """(H2O)32 / 6-31G** (800 functions, direct J/K), where the FP64 atomics' order has moved converged energies
    by millihartrees between runs: two deterministic runs give the same bits."""
    _print_case("water32_631gss")
    kw = dict(alpha=0.5, init_guess=0.0, normal_iter_after_diis=0, energy_tolerance=1e-8, density_tolerance=1e-6,
              diis_tolerance=1e-6, max_iterations=300)
    out = []
    for _ in range(2):
        eng = gpu_engine_factory(schwarz_threshold=1e-12, exact_transfer=1, deterministic=1).set_molecule(
            basis("6-31g_st__st_.nw"), g.water32_cfg3())
        r = eng.scf(incore=0, **kw)
        print("water32 deterministic: E %.12f iterations %d converged %d" % (r.energy, r.iterations, r.converged),
              flush=True)
        out.append((r.energy, r.iterations, r.converged))
        eng.close()
    assert out[0] == out[1]
    assert out[0][2]
