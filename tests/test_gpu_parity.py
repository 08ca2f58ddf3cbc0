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


MIXED_CASES = [
    ("ethane_631gs", "6-31g_st_.1.nw", g.alkane_zigzag(2, seed=5)),
    ("water4_631gss", "6-31g_st__st_.nw", g.water_cluster(4, seed=2024)),
]
MIXED_KW = dict(alpha=0.5, init_guess=0.0, normal_iter_after_diis=0)


@pytest.mark.parametrize("name,bfile,atoms", MIXED_CASES, ids=[c[0] for c in MIXED_CASES])
def test_mixed_jk_close_to_fp64(gpu_engine_factory, name, bfile, atoms):
    """Mixed-precision J/K (quartets with Schwarz bound below 1e-6 in FP32) against the FP64
    build on the converged density: the FP32 band's integrals are below 1e-6 with errors below
    1e-3 of that, so J and K differ by far less than 1e-7."""
    _print_case(name)
    e64 = gpu_engine_factory().set_molecule(basis(bfile), atoms)
    e32 = gpu_engine_factory(precision=1, fp32_threshold=1e-6).set_molecule(basis(bfile), atoms)
    _, D = e64.scf(want_density=True, **MIXED_KW)
    D = D[0]
    J64, K64 = e64.build_jk(D)
    e32.reset_kernel_stats()
    J32, K32 = e32.build_jk(D)
    dj, dk = np.abs(J32 - J64).max(), np.abs(K32 - K64).max()
    print("max|dJ| %.3e max|dK| %.3e" % (dj, dk), flush=True)
    fp32_quartets = sum(s.quartets for s in e32.kernel_stats() if s.fp32)
    fp64_quartets = sum(s.quartets for s in e32.kernel_stats() if not s.fp32)
    print("fp32 quartets %d fp64 quartets %d" % (fp32_quartets, fp64_quartets), flush=True)
    assert fp32_quartets > 0
    assert dj <= 1e-7 and dk <= 1e-7, (dj, dk)


MIXED_ORACLE_CASES = [
    ("ethane_631gs", "6-31g_st_.1.nw", g.alkane_zigzag(2, seed=5)),
    ("water2_631gss", "6-31g_st__st_.nw", g.water_cluster(2, seed=2024)),
]


@pytest.mark.parametrize("name,bfile,atoms", MIXED_ORACLE_CASES, ids=[c[0] for c in MIXED_ORACLE_CASES])
def test_mixed_tier_matches_oracle(gpu_engine_factory, name, bfile, atoms):
    """The mixed-precision tier against the oracle rather than the FP64 GPU path: J/K of the oracle's converged
    density within 1e-7 of the oracle's N^4 loop over its own ERIs, and the mixed SCF energy within 1e-6 Eh
    (the tier's stated bound) of the oracle's SCF."""
    _print_case(name)
    mol = refcpu.Molecule(basis(bfile), atoms)
    ref, Dref = mol.scf(**MIXED_KW)
    eri = mol.eri_packed()
    n = mol.nbf
    Jr, Kr = refcpu.jk_from_packed(n, eri, Dref[0])
    e32 = gpu_engine_factory(precision=1, fp32_threshold=1e-6).set_molecule(basis(bfile), atoms)
    e32.reset_kernel_stats()
    J, K = e32.build_jk(Dref[0])
    assert sum(s.quartets for s in e32.kernel_stats() if s.fp32) > 0
    dj, dk = np.abs(J - Jr).max(), np.abs(K - Kr).max()
    r = e32.scf(**MIXED_KW)
    print("mixed vs oracle: max|dJ| %.2e max|dK| %.2e  E %.12f oracle %.12f" % (dj, dk, r.energy, ref.energy),
          flush=True)
    assert dj <= 1e-7 and dk <= 1e-7
    assert r.converged and r.fp32_builds > 0 and abs(r.energy - ref.energy) <= 1e-6


@pytest.mark.parametrize("name,bfile,atoms", MIXED_CASES, ids=[c[0] for c in MIXED_CASES])
def test_mixed_scf_energy_matches_fp64(gpu_engine_factory, name, bfile, atoms):
    """Mixed-precision SCF: FP32 band until the criteria are nearly met, then all-FP64 steps (at
    least 3) before convergence is accepted.  The energy must agree with the all-FP64 run to
    1e-6 Eh (the tier's target); in practice the final FP64 steps bring it to the FP64 result."""
    _print_case(name)
    e64 = gpu_engine_factory().set_molecule(basis(bfile), atoms)
    e32 = gpu_engine_factory(precision=1, fp32_threshold=1e-6).set_molecule(basis(bfile), atoms)
    r64 = e64.scf(**MIXED_KW)
    r32 = e32.scf(**MIXED_KW)
    print("E64 %.12f E32 %.12f dE %.3e iters %d/%d fp32 builds %d" % (
        r64.energy, r32.energy, r32.energy - r64.energy, r64.iterations, r32.iterations, r32.fp32_builds), flush=True)
    assert r64.converged and r32.converged
    assert r32.fp32_builds > 0
    assert abs(r32.energy - r64.energy) <= 1e-6


SERVE_MOLS = [("water1", g.water_cluster(1, seed=11)), ("ethane", g.alkane_zigzag(2, seed=12)),
              ("water3", g.water_cluster(3, seed=13)), ("butane", g.alkane_zigzag(4, seed=14)),
              ("water2", g.water_cluster(2, seed=15)), ("methane", g.alkane_zigzag(1, seed=16))]
SERVE_SCF = dict(alpha=0.5, init_guess=0.0, normal_iter_after_diis=0, energy_tolerance=1e-8,
                 density_tolerance=1e-6, diis_tolerance=1e-6)


def _serve_all(server, bfile, mols):
    ids = {}
    for i, (name, atoms) in enumerate(mols):
        ids[server.submit(bfile, atoms, tag=i)] = name
    out = {}
    while len(out) < len(mols):
        for r in server.poll(16, 2000):
            out[ids[r.id]] = r
    return out


@pytest.mark.parametrize("incore", [0, 1])
def test_serving_results_match_alone(gpu_engine_factory, incore):
    """Continuous batching: six molecules submitted at once to three slots run concurrently; each
    energy equals the molecule's SCF run alone (1e-9 Eh: only the order of FP64 atomics differs),
    with direct J/K and with each request's in-core ERI store."""
    import hartreefock_amd as hf
    bfile = basis("6-31g.1.nw")
    alone = {}
    for name, atoms in SERVE_MOLS:
        _print_case(name)
        alone[name] = gpu_engine_factory().set_molecule(bfile, atoms).scf(**SERVE_SCF).energy
    srv = hf.Server(slots=3, scf=dict(SERVE_SCF, incore=incore))
    try:
        got = _serve_all(srv, bfile, SERVE_MOLS)
    finally:
        srv.close()
    overlap = 0
    for name, _ in SERVE_MOLS:
        r = got[name]
        print("%-8s alone %.10f served %.10f  arrive %.1f admit %.1f done %.1f" % (
            name, alone[name], r.energy, r.arrival_ms, r.admit_ms, r.done_ms), flush=True)
        assert r.status == 0 and r.converged
        assert abs(r.energy - alone[name]) <= 1e-9
        overlap += sum(1 for o in got.values() if o is not r and o.admit_ms < r.done_ms and r.admit_ms < o.done_ms)
    assert overlap > 0            # requests did run concurrently


def test_serving_admission_by_footprint(gpu_engine_factory):
    """With an HBM budget that holds one request at a time, requests run one after another (no two
    admitted intervals overlap), first come first served, and the results are unchanged."""
    import hartreefock_amd as hf
    bfile = basis("6-31g.1.nw")
    mols = SERVE_MOLS[:4]
    budget = max(hf.serve_footprint(gpu_engine_factory().set_molecule(bfile, a).nbf) for _, a in mols) * 1.5
    srv = hf.Server(slots=3, hbm_budget=budget, scf=SERVE_SCF)
    try:
        got = _serve_all(srv, bfile, mols)
    finally:
        srv.close()
    rs = sorted(got.values(), key=lambda r: r.admit_ms)
    for a, b in zip(rs, rs[1:]):
        assert b.admit_ms >= a.done_ms - 1e-6
        assert b.id > a.id
    for name, atoms in mols:
        e = gpu_engine_factory().set_molecule(bfile, atoms).scf(**SERVE_SCF).energy
        assert abs(got[name].energy - e) <= 1e-9


# ---- at scale: many blocks per class, blocks spanning several bra pairs, primitive slices ----
SCALE_CASES = [
    ("pentane_631gs", "6-31g_st_.1.nw", g.alkane_zigzag(5, seed=31)),       # 99 functions, d shells
    ("water6_631gss", "6-31g_st__st_.nw", g.water_cluster(6, seed=32)),     # 150 functions, p on H
    ("pentyl_631pgs", "6-31+g_st_.1.nw", g.alkane_zigzag(5, seed=33, drop_last_h=True)),   # 117, diffuse sp on C
]


@pytest.mark.parametrize("name,bfile,atoms", SCALE_CASES, ids=[c[0] for c in SCALE_CASES])
def test_jk_at_scale_matches_packed_contraction(gpu_engine_factory, name, bfile, atoms):
    """Direct J/K (screened task lists, LDS exchange slots, primitive slices) against the oracle's
    J/K loop over the packed ERIs of the same molecule.  The packed store takes whole quartets
    from a separate unscreened task list, and is itself checked against the oracle's ERIs on the
    small molecules above; here it stands in for the oracle's ERIs, which take minutes at this size."""
    _print_case(name)
    eng = gpu_engine_factory(schwarz_threshold=0.0).set_molecule(basis(bfile), atoms)
    n = eng.nbf
    eri = eng.eri_packed()
    rng = np.random.default_rng(11)
    X = rng.normal(size=(n, n)) * 0.1
    D = X + X.T
    J, K = eng.build_jk(D)
    Jref, Kref = refcpu.jk_from_packed(n, eri, D)
    assert np.abs(J - Jref).max() <= 1e-11 * max(1.0, np.abs(Jref).max())
    assert np.abs(K - Kref).max() <= 1e-11 * max(1.0, np.abs(Kref).max())
    # the production threshold changes J/K only by the screened-out quartets
    eng12 = gpu_engine_factory(schwarz_threshold=1e-12).set_molecule(basis(bfile), atoms)
    J12, K12 = eng12.build_jk(D)
    assert np.abs(J12 - Jref).max() <= 1e-8 and np.abs(K12 - Kref).max() <= 1e-8


DENSITY_CASES = [
    ("water2_631gs_local_sp2", "6-31g_st_.1.nw", g.water_cluster(2, seed=41), False),   # n <= 96: one workgroup
    ("pentane_631gs_fused_sp2", "6-31g_st_.1.nw", g.alkane_zigzag(5, seed=42), False),  # n = 99: one kernel per step
    ("octane_631gs_fused_sp2", "6-31g_st_.1.nw", g.alkane_zigzag(8, seed=42), False),   # n = 156: one kernel per step
    ("ch3_631gs_uhf", "6-31g_st_.1.nw", [(6, (0.0, 0.0, 0.0)), (1, (2.04, 0.0, 0.0)), (1, (-1.02, 1.767, 0.0)),
                                         (1, (-1.02, -1.767, 0.0))], True),
]


@pytest.mark.parametrize("name,bfile,atoms,uhf", DENSITY_CASES, ids=[c[0] for c in DENSITY_CASES])
def test_sp2_density_matches_diagonalisation(gpu_engine_factory, name, bfile, atoms, uhf):
    """SP2 purification (both device paths) against diagonalising every step, the reference's route:
    the same converged energy within 1e-8 Eh."""
    _print_case(name)
    kw = dict(alpha=0.5, init_guess=0.0, normal_iter_after_diis=0, energy_tolerance=1e-10,
              density_tolerance=1e-8, diis_tolerance=1e-8, max_iterations=200)
    eng = gpu_engine_factory(schwarz_threshold=1e-12, exact_transfer=1).set_molecule(basis(bfile), atoms)
    r_sp2 = eng.scf(unrestricted=uhf, density_method=0, **kw)
    r_diag = eng.scf(unrestricted=uhf, density_method=1, **kw)
    assert r_sp2.converged and r_diag.converged
    assert abs(r_sp2.energy - r_diag.energy) <= 1e-8
    assert r_sp2.sp2_steps > 0


# ---- in-core J/K: ERIs evaluated once into the pair matrix, contracted every build ----
@pytest.mark.parametrize("name,bfile,atoms", SMALL[3:6], ids=[s[0] for s in SMALL[3:6]])
@pytest.mark.parametrize("ndens", [1, 2])
def test_incore_jk_matches_oracle(gpu_engine_factory, name, bfile, atoms, ndens):
    """In-core J/K (pair-matrix store, one workgroup per bra pair) against the oracle's J/K loop over
    its own packed ERIs (the reference's InitFockMatrix contraction), 1e-10 relative."""
    _print_case(name)
    eng = gpu_engine_factory(schwarz_threshold=0.0).set_molecule(basis(bfile), atoms)
    mol = refcpu.Molecule(basis(bfile), atoms)
    n = eng.nbf
    eri = mol.eri_packed()
    rng = np.random.default_rng(17)
    Ds = []
    for _ in range(ndens):
        X = rng.normal(size=(n, n))
        Ds.append(0.5 * (X + X.T))
    D = np.array(Ds) if ndens == 2 else Ds[0]
    J, K = eng.build_jk_incore(D)
    Jref = sum(refcpu.jk_from_packed(n, eri, d)[0] for d in Ds)
    assert np.abs(J - Jref).max() <= 1e-10 * max(1.0, np.abs(Jref).max())
    Kg = K if ndens == 2 else [K]
    for d, k in zip(Ds, Kg):
        Kref = refcpu.jk_from_packed(n, eri, d)[1]
        assert np.abs(k - Kref).max() <= 1e-10 * max(1.0, np.abs(Kref).max())


@pytest.mark.parametrize("name,bfile,atoms", SCALE_CASES, ids=[c[0] for c in SCALE_CASES])
def test_incore_jk_at_scale_matches_packed_contraction(gpu_engine_factory, name, bfile, atoms):
    """In-core J/K at 99 and 150 functions (two and three column registers per lane, rows longer than
    a wavefront) against the oracle's J/K loop over the GPU's packed store, and against direct J/K."""
    _print_case(name)
    eng = gpu_engine_factory(schwarz_threshold=0.0).set_molecule(basis(bfile), atoms)
    n = eng.nbf
    eri = eng.eri_packed()
    rng = np.random.default_rng(19)
    X = rng.normal(size=(2, n, n)) * 0.1
    D = X + X.transpose(0, 2, 1)
    J, K = eng.build_jk_incore(D)
    Jref = sum(refcpu.jk_from_packed(n, eri, d)[0] for d in D)
    assert np.abs(J - Jref).max() <= 1e-11 * max(1.0, np.abs(Jref).max())
    for d, k in zip(D, K):
        Kref = refcpu.jk_from_packed(n, eri, d)[1]
        assert np.abs(k - Kref).max() <= 1e-11 * max(1.0, np.abs(Kref).max())
    Jd, Kd = eng.build_jk(D)
    assert np.abs(J - Jd).max() <= 1e-11 * max(1.0, np.abs(Jd).max())
    assert np.abs(K - Kd).max() <= 1e-11 * max(1.0, np.abs(Kd).max())


NSHAPES = 8        # tile shapes of the in-core contraction (hf_options.incore_shape 0..7)
SHAPE_CASES = [
    ("water4_631g", "6-31g.1.nw", g.water_cluster(4, seed=41)),            # 52 functions: one 64-column chunk
    ("pentane_631gs", "6-31g_st_.1.nw", g.alkane_zigzag(5, seed=31)),       # 99: two chunks of 64, one of 128
    ("water6_631gss", "6-31g_st__st_.nw", g.water_cluster(6, seed=32)),     # 150: three of 64, two of 128
    ("water24_631g", "6-31g.1.nw", g.water_cluster(24, seed=42)),           # 312: five of 64, three of 128
]


@pytest.mark.parametrize("name,bfile,atoms", SHAPE_CASES, ids=[c[0] for c in SHAPE_CASES])
def test_incore_every_shape_matches_direct(gpu_engine_factory, name, bfile, atoms):
    """Every tile shape of the in-core contraction forced in turn (one or two columns per lane, one or two
    rows per group, two register buffers in turn, one or three store rows per block), for one and two densities, on one pair-matrix store: against the direct J/K
    build of the same densities at 1e-11 relative, and the shapes against each other; the autotuner may
    pick any of them for a size."""
    import hartreefock_amd as hf
    from devmem import DevArray, sync
    _print_case(name)
    ref = gpu_engine_factory(schwarz_threshold=0.0).set_molecule(basis(bfile), atoms)
    n = ref.nbf
    lg = hf.lib().hf_pair_matrix_len(n)
    dG = DevArray(lg, fill=0.0)
    ref.eri_pairs_device(dG.ptr, lg)
    rng = np.random.default_rng(23 + n)
    X = rng.normal(size=(2, n, n)) * 0.1
    D2 = X + X.transpose(0, 2, 1)
    cases = []
    for nd in (1, 2):
        D = D2[:nd]
        Jd, Kd = ref.build_jk(D if nd == 2 else D[0])
        cases.append((nd, DevArray.of(D), Jd, Kd.reshape(nd, n, n)))
    sync()
    first = {}
    for shape in range(NSHAPES):
        eng = gpu_engine_factory(schwarz_threshold=0.0, incore_shape=shape).set_molecule(basis(bfile), atoms)
        for nd, dD, Jd, Kd in cases:
            dJ, dK = DevArray((n, n), fill=0.0), DevArray((nd, n, n), fill=0.0)
            eng.build_jk_incore_device(dG.ptr, dD.ptr, dJ.ptr, dK.ptr, ndens=nd)
            sync()
            J, K = dJ.get(), dK.get()
            dJ.free(); dK.free()
            ej = np.abs(J - Jd).max() / max(1.0, np.abs(Jd).max())
            ek = np.abs(K - Kd).max() / max(1.0, np.abs(Kd).max())
            print("n %d shape %d ndens %d: rel |dJ| %.2e |dK| %.2e" % (n, shape, nd, ej, ek), flush=True)
            assert ej <= 1e-11 and ek <= 1e-11, (shape, nd)
            if nd in first:
                J0, K0 = first[nd]
                assert np.abs(J - J0).max() <= 1e-12 * max(1.0, np.abs(J0).max()), (shape, nd)
                assert np.abs(K - K0).max() <= 1e-12 * max(1.0, np.abs(K0).max()), (shape, nd)
            else:
                first[nd] = (J, K)
        eng.close()
    # the shape count is pinned: one past the last shape is refused
    eng = gpu_engine_factory(schwarz_threshold=0.0, incore_shape=NSHAPES).set_molecule(basis(bfile), atoms)
    dJ, dK = DevArray((n, n), fill=0.0), DevArray((1, n, n), fill=0.0)
    with pytest.raises(hf.HFError):
        eng.build_jk_incore_device(dG.ptr, cases[0][1].ptr, dJ.ptr, dK.ptr, ndens=1)
    eng.close()
    for c in cases:
        c[1].free()
    dG.free()
    ref.close()


INCORE_SCF_CASES = [
    ("h2o_crawford_sto3g", "sto3g.txt", g.water_crawford(), False, dict(alpha=0.5, init_guess=0.7)),
    ("ethane_631gs", "6-31g_st_.1.nw", g.alkane_zigzag(2, seed=5), False, dict(alpha=0.5, init_guess=0.0)),
    ("ch3_631gs_uhf", "6-31g_st_.1.nw", UHF_CASES[1][2], True, dict(alpha=0.5, init_guess=0.0)),
]


@pytest.mark.parametrize("name,bfile,atoms,uhf,kw", INCORE_SCF_CASES, ids=[c[0] for c in INCORE_SCF_CASES])
def test_incore_scf_matches_oracle(gpu_engine_factory, name, bfile, atoms, uhf, kw):
    """SCF with in-core J/K: the oracle's energy within 1e-8 Eh (RHF and UHF)."""
    _print_case(name)
    kw = dict(kw, normal_iter_after_diis=0)
    eng = gpu_engine_factory(schwarz_threshold=0.0).set_molecule(basis(bfile), atoms)
    r = eng.scf(unrestricted=uhf, incore=1, **kw)
    assert r.incore == 1
    ref, _ = refcpu.Molecule(basis(bfile), atoms).scf(unrestricted=uhf, **kw)
    assert r.converged and ref.converged
    assert abs(r.energy - ref.energy) <= E_TOL, (r.energy, ref.energy)


@pytest.mark.parametrize("name,bfile,atoms", MIXED_CASES, ids=[c[0] for c in MIXED_CASES])
def test_incore_mixed_scf_matches_fp64(gpu_engine_factory, name, bfile, atoms):
    """Mixed tier with in-core J/K (FP32 copy of the store until the final FP64 steps) against the
    all-FP64 direct SCF: |dE| <= 1e-6 Eh; and the FP64 in-core SCF against the direct one, 1e-8 Eh."""
    _print_case(name)
    kw = dict(MIXED_KW, energy_tolerance=1e-8, density_tolerance=1e-6, diis_tolerance=1e-6)
    e64 = gpu_engine_factory(schwarz_threshold=1e-12, exact_transfer=1).set_molecule(basis(bfile), atoms).scf(**kw)
    inc = gpu_engine_factory(schwarz_threshold=1e-12, exact_transfer=1).set_molecule(basis(bfile), atoms)
    r64 = inc.scf(incore=1, **kw)
    mix = gpu_engine_factory(schwarz_threshold=1e-12, exact_transfer=1, precision=1).set_molecule(basis(bfile), atoms)
    rm = mix.scf(incore=1, **kw)
    print("iterations: FP64 in-core %d, mixed in-core %d (%d builds from the FP32 copy)"
          % (r64.iterations, rm.iterations, rm.fp32_builds), flush=True)
    assert e64.converged and r64.converged and rm.converged
    assert r64.incore == 1 and rm.incore == 1 and rm.fp32_builds > 0
    assert abs(r64.energy - e64.energy) <= 1e-8
    assert abs(rm.energy - e64.energy) <= 1e-6
    # outside the FP32 increments every mixed in-core build is a full FP64 build and one FP64 step ends the SCF, so
    # the tier costs no extra iterations beyond the FP32 rounding's small effect on the trajectory
    assert rm.iterations <= r64.iterations + 2


# ---- after the SCF: properties and MP2 ----
TIGHT = dict(alpha=0.5, init_guess=0.0, normal_iter_after_diis=0, energy_tolerance=1e-11, density_tolerance=1e-9,
             diis_tolerance=1e-9)
POST_CASES = [
    ("h2o_crawford_sto3g", "sto3g.txt", g.water_crawford(), False),
    ("ethane_631gs", "6-31g_st_.1.nw", g.alkane_zigzag(2, seed=5), False),
    ("ch3_631gs_uhf", "6-31g_st_.1.nw", UHF_CASES[1][2], True),
]


@pytest.mark.parametrize("name,bfile,atoms,uhf", POST_CASES, ids=[c[0] for c in POST_CASES])
def test_properties_match_oracle(gpu_engine_factory, name, bfile, atoms, uhf):
    """Mulliken charges and dipole (device overlap and dipole integrals) against the oracle's on the
    same density, 1e-10; H2O/STO-3G against Crawford's known answers."""
    _print_case(name)
    eng = gpu_engine_factory(schwarz_threshold=0.0).set_molecule(basis(bfile), atoms)
    r, D = eng.scf(unrestricted=uhf, want_density=True, **TIGHT)
    assert r.converged
    q, mu = eng.properties(D if uhf else D[0])
    mol = refcpu.Molecule(basis(bfile), atoms)
    S = mol.one_electron()[0]
    Dt = D.sum(0)
    Z = [a[0] for a in atoms]
    R = [a[1] for a in atoms]
    q_ref = refcpu.mulliken_charges(Dt, S, Z, mol.bf_info()[2])
    mu_ref = refcpu.dipole(Dt, mol.moment(), Z, R)
    assert np.abs(q - q_ref).max() <= 1e-10 and np.abs(mu - mu_ref).max() <= 1e-10
    assert abs(q.sum() - (sum(Z) - sum(eng.electrons()))) <= 1e-10
    if name == "h2o_crawford_sto3g":
        assert np.abs(q - np.array([-0.253146052405, 0.126573026202, 0.126573026202])).max() <= 1e-8
        assert abs(mu[1] - 0.603521296525) <= 1e-8 and abs(mu[0]) <= 1e-10 and abs(mu[2]) <= 1e-10


@pytest.mark.parametrize("name,bfile,atoms,uhf", POST_CASES, ids=[c[0] for c in POST_CASES])
@pytest.mark.parametrize("incore", [0, 1])
def test_mp2_matches_oracle(gpu_engine_factory, name, bfile, atoms, uhf, incore):
    """MP2 correlation energy (two FP64 GEMMs over the in-core store) against the oracle's restatement
    of the reference's CalculateMp2Energy on the oracle's own SCF, 1e-8 Eh; H2O/STO-3G against Crawford's
    -0.049149636120 Eh.  UHF follows the reference's formula as written (same-spin sums times 1/4)."""
    _print_case(name)
    eng = gpu_engine_factory(schwarz_threshold=0.0).set_molecule(basis(bfile), atoms)
    r = eng.scf(unrestricted=uhf, mp2=1, incore=incore, **TIGHT)
    assert r.converged and r.mp2_energy < 0
    mol = refcpu.Molecule(basis(bfile), atoms)
    ref, D = mol.scf(unrestricted=uhf, alpha=0.5, init_guess=0.0, normal_iter_after_diis=0)
    assert ref.converged
    n = mol.nbf
    S, T, V = mol.one_electron()
    eri = mol.eri_packed()
    na, nb = mol.electrons()
    if uhf:
        e2 = refcpu.mp2_uhf_reference(n, eri, D[0], D[1], T + V, S, na, nb)
    else:
        e2 = refcpu.mp2_rhf(n, eri, D[0], T + V, S, na)
    print("mp2 gpu %.12f oracle %.12f" % (r.mp2_energy, e2), flush=True)
    assert abs(r.mp2_energy - e2) <= 1e-8
    if name == "h2o_crawford_sto3g":
        assert abs(r.mp2_energy - (-0.049149636120)) <= 1e-8


def test_cpp_dropin_runs(tmp_path):
    """The reference-style C++ callers (tests/cpp/dropin_example.cpp, over include/hartreefock_amd.hpp and the C ABI)
    against the oracle: atoms taken from Chemistry::Basis::atoms, RHF Init/Calculate; the public h; ERIs by index and
    by orbital pointers; getBoysFunctions; the SCF driven by Step(iter); an H2 bond scan moving atoms with
    SetCenterForShells; LoadXYZ with a UHF cation in an electric field."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "dropin_example"
    cmd = ["g++", "-std=c++17", "-O1", "-I" + os.path.join(root, "include"),
           os.path.join(root, "tests", "cpp", "dropin_example.cpp"),
           "-L" + os.path.join(root, "hartreefock_amd"), "-lhfgpu", "-o", str(exe)]
    assert subprocess.run(cmd, capture_output=True, text=True).returncode == 0
    env = dict(os.environ, LD_LIBRARY_PATH=os.path.join(root, "hartreefock_amd") + ":" +
               os.environ.get("LD_LIBRARY_PATH", ""))
    p = subprocess.run([str(exe), basis("sto3g.txt"), str(tmp_path / "water.xyz")], capture_output=True, text=True,
                       env=env, timeout=100)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0
    vals = dict(line.split(" = ", 1) for line in p.stdout.splitlines() if " = " in line)
    kw = dict(alpha=0.5, init_guess=0.0, normal_iter_after_diis=0)
    wat = [(1, (0.0, 1.43233673, -0.96104039)), (1, (0.0, -1.43233673, -0.96104039)), (8, (0.0, 0.0, 0.24026010))]
    mol = refcpu.Molecule(basis("sto3g.txt"), wat)
    ref, _ = mol.scf(**kw)
    assert int(vals["nbf"]) == 7 and vals["electrons"].split() == ["5", "5"]
    assert abs(float(vals["E"]) - ref.energy) <= 1e-8 and vals["converged"] == "1"
    assert abs(float(vals["E_steps"]) - ref.energy) <= 1e-8
    S, T, V = mol.one_electron()
    assert abs(float(vals["h00"]) - (T + V)[0, 0]) <= 1e-12
    refcpu.set_exact_transfer(False)
    e = mol.eri(2, 3, 5, 6)
    assert abs(float(vals["eri_int"]) - e) <= 1e-12 and abs(float(vals["eri_orb"]) - e) <= 1e-12
    assert int(vals["boys_n"]) == 6 and abs(float(vals["boys0"]) - refcpu.boys(5, 2.5)[0]) <= 1e-14
    assert float(vals["E(MP2)"]) < 0
    for r in (1.2, 1.4, 1.6):
        er, conv, _ = vals["scan %.1f" % r].split()
        h2, _ = refcpu.Molecule(basis("sto3g.txt"), [(1, (-r / 2, 0.0, 0.0)), (1, (r / 2, 0.0, 0.0))]).scf(
            use_diis=0, **kw)
        assert conv == "1" and abs(float(er) - h2.energy) <= 1e-8
    cat = refcpu.Molecule(basis("sto3g.txt"), g.water_cfg1()).set_electrons(5, 4).set_field([0.0, 0.0, 0.01])
    u, _ = cat.scf(unrestricted=True, **kw)
    assert vals["loaded"] == "1" and vals["uhf_converged"] == "1"
    assert abs(float(vals["E_uhf_field"]) - u.energy) <= 1e-8
    assert abs(float(vals["E_field_nuc"]) - (-sum(z * x[2] * 0.01 for z, x in g.water_cfg1()))) <= 1e-13


@pytest.mark.parametrize("prec,incr", [(0, 1), (1, 0)], ids=["fp64_incremental", "mixed"])
def test_incore_uhf_density_changes_match_direct(gpu_engine_factory, prec, incr):
    """UHF with the in-core store and J/K of density changes (FP64 incremental builds, and the mixed
    tier's FP32 store on density changes): the direct FP64 energy within 1e-8 Eh.  This SCF converges
    slowly (~190 iterations): with criteria of 1e-8 Eh / 1e-6 its stopping point scatters by ~1e-6 Eh from
    run to run (a plain FP64 in-core SCF and the direct one stopped 1.2e-6 Eh apart), so both SCFs are
    converged tightly and land on the same stationary point (measured agreement 2e-11 Eh)."""
    atoms = UHF_CASES[1][2]
    _print_case("ch3_631gs_uhf prec=%d incremental=%d" % (prec, incr))
    kw = dict(alpha=0.5, init_guess=0.0, normal_iter_after_diis=0, energy_tolerance=1e-10, density_tolerance=1e-8,
              diis_tolerance=1e-8, max_iterations=600)
    ref = gpu_engine_factory(schwarz_threshold=1e-12, exact_transfer=1).set_molecule(
        basis("6-31g_st_.1.nw"), atoms).scf(unrestricted=True, **kw)
    eng = gpu_engine_factory(schwarz_threshold=1e-12, exact_transfer=1, precision=prec).set_molecule(
        basis("6-31g_st_.1.nw"), atoms)
    r = eng.scf(unrestricted=True, incore=1, incremental_fock=incr, **kw)
    print("uhf direct %.10f in-core %.10f s2 %.4f %.4f" % (ref.energy, r.energy, ref.s2, r.s2), flush=True)
    assert ref.converged and r.converged and r.incore == 1
    assert abs(r.energy - ref.energy) <= 1e-8 and abs(r.s2 - ref.s2) <= 1e-6


def test_mp2_at_scale_matches_oracle_restatement(gpu_engine_factory):
    """MP2 at 99 functions (d shells, 21 occupied and 78 virtual orbitals): the GPU's GEMM transform over the
    in-core store against the oracle's restatement of the reference's MP2 applied to the GPU's own packed ERIs
    and converged density (the oracle's ERIs take minutes at this size), 1e-8 Eh."""
    _print_case("pentane_631gs mp2")
    atoms = g.alkane_zigzag(5, seed=31)
    eng = gpu_engine_factory(schwarz_threshold=0.0, exact_transfer=1).set_molecule(basis("6-31g_st_.1.nw"), atoms)
    r, D = eng.scf(mp2=1, incore=1, want_density=True, **TIGHT)
    assert r.converged and r.incore == 1
    n = eng.nbf
    S, T, V = eng.one_electron()
    e2 = refcpu.mp2_rhf(n, eng.eri_packed(), D[0], T + V, S, eng.electrons()[0])
    print("mp2 gpu %.12f oracle %.12f" % (r.mp2_energy, e2), flush=True)
    assert abs(r.mp2_energy - e2) <= 1e-8
