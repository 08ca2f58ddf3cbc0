"""Lockstep SCF of several molecules (hf_scf_batch): each phase of an iteration is one launch over every molecule
still iterating.  A molecule's result must not depend on the batch it ran in: against the same molecule's hf_scf
alone, energies agree to the FP64 atomics' rounding (1e-9 Eh) and, in deterministic mode, bit for bit with the
same iteration count; against the oracle to 1e-8 Eh.  Failures (an iteration cap, a non-finite SCF) stay with
their molecule."""
import numpy as np
import pytest

from conftest import basis
from hartreefock_amd import geometry as g
from oracle import refcpu

pytestmark = pytest.mark.gpu

E_TOL = 1e-8
KW = dict(alpha=0.5, init_guess=0.0, normal_iter_after_diis=0)

OH = [(8, (0.0, 0.0, 0.0)), (1, (0.0, 0.0, 1.83))]
CH3 = [(6, (0.0, 0.0, 0.0)), (1, (2.04, 0.0, 0.0)), (1, (-1.02, 1.767, 0.0)), (1, (-1.02, -1.767, 0.0))]

# sizes on both sides of the single-workgroup SP2 limit (n <= 96) and RHF beside UHF
MIX = [
    ("h2o_sto3g", "sto3g.txt", g.water_cfg1(), False),
    ("oh_sto3g_uhf", "sto3g.txt", OH, True),
    ("ethane_631gs", "6-31g_st_.1.nw", g.alkane_zigzag(2, seed=5), False),
    ("ch3_631gs_uhf", "6-31g_st_.1.nw", CH3, True),
    ("water4_631gss", "6-31g_st__st_.nw", g.water_cluster(4, seed=7), False),    # 100 functions: stepped SP2
    ("pentane_631gs", "6-31g_st_.1.nw", g.alkane_zigzag(5, seed=31), False),     # 99 functions
    ("octane_631gs", "6-31g_st_.1.nw", g.alkane_zigzag(8, seed=13), False),      # 156 functions
]


def _engines(factory, mols, **opts):
    return [factory(**opts).set_molecule(basis(b), a) for _, b, a, _ in mols]


@pytest.mark.parametrize("incore", [0, 1], ids=["direct", "incore"])
@pytest.mark.parametrize("det", [0, 1], ids=["atomic", "deterministic"])
def test_batch_matches_alone(gpu_engine_factory, incore, det):
    import hartreefock_amd as hf
    print("case batch incore=%d deterministic=%d" % (incore, det), flush=True)
    opts = dict(deterministic=det, exact_transfer=1)
    engs = _engines(gpu_engine_factory, MIX, **opts)
    out = hf.scf_batch(engs, unrestricted=[u for *_, u in MIX], incore=incore, **KW)
    for (name, b, atoms, uhf), (st, r) in zip(MIX, out):
        a = gpu_engine_factory(**opts).set_molecule(basis(b), atoms).scf(unrestricted=uhf, incore=incore, **KW)
        print("%-16s batch %.12f (%d it)  alone %.12f (%d it)" % (name, r.energy, r.iterations, a.energy,
                                                                   a.iterations), flush=True)
        assert st == 0 and r.converged and a.converged
        if det:
            assert r.energy == a.energy and r.iterations == a.iterations
        else:
            assert abs(r.energy - a.energy) <= 1e-9
        if uhf:
            assert abs(r.s2 - a.s2) <= 1e-9


ORACLE = [MIX[0], MIX[1], MIX[2]]


def test_batch_matches_oracle(gpu_engine_factory):
    import hartreefock_amd as hf
    engs = _engines(gpu_engine_factory, ORACLE, schwarz_threshold=0.0)
    out = hf.scf_batch(engs, unrestricted=[u for *_, u in ORACLE], **KW)
    for (name, b, atoms, uhf), (st, r) in zip(ORACLE, out):
        ref, _ = refcpu.Molecule(basis(b), atoms).scf(unrestricted=uhf, **KW)
        print("%-16s batch %.12f oracle %.12f" % (name, r.energy, ref.energy), flush=True)
        assert st == 0 and r.converged and ref.converged
        assert abs(r.energy - ref.energy) <= E_TOL


def test_batch_failures_stay_with_their_molecule(gpu_engine_factory):
    """One molecule capped at 2 iterations (HF_E_NOT_CONVERGED), one with coincident nuclei (HF_E_NUMERIC at its
    first iteration): the others converge to their solo energies."""
    import hartreefock_amd as hf
    bad = [(8, (0.0, 0.0, 0.0)), (1, (0.0, 0.0, 0.0)), (1, (0.0, 1.4, 1.0))]
    mols = [MIX[2], ("coincident", "sto3g.txt", bad, False), MIX[0], ("capped", "6-31g_st_.1.nw",
                                                                    g.alkane_zigzag(3, seed=5), False)]
    engs = _engines(gpu_engine_factory, mols)
    params = [dict(KW), dict(KW), dict(KW), dict(KW, max_iterations=2)]
    import ctypes
    P = (hf.ScfParams * len(mols))()
    for i, kw in enumerate(params):
        hf.lib().hf_scf_defaults(ctypes.byref(P[i]))
        for k, v in kw.items():
            setattr(P[i], k, v)
    hs = (ctypes.c_void_p * len(mols))(*[e.h for e in engs])
    res = (hf.ScfResult * len(mols))()
    st = (ctypes.c_int * len(mols))()
    rc = hf.lib().hf_scf_batch(hs, len(mols), P, res, st)
    print("statuses", list(st), "rc", rc, flush=True)
    assert st[1] == -9 and st[3] == -4 and rc == -9
    assert res[3].iterations == 2 and res[3].converged == 0
    for i in (0, 2):
        name, b, atoms, uhf = mols[i]
        a = gpu_engine_factory().set_molecule(basis(b), atoms).scf(**KW)
        assert st[i] == 0 and res[i].converged
        assert abs(res[i].energy - a.energy) <= 1e-9


def test_batch_rejects_a_context_twice(gpu_engine_factory):
    import ctypes
    import hartreefock_amd as hf
    e = gpu_engine_factory().set_molecule(basis("sto3g.txt"), g.water_cfg1())
    P = (hf.ScfParams * 2)()
    for i in range(2):
        hf.lib().hf_scf_defaults(ctypes.byref(P[i]))
    hs = (ctypes.c_void_p * 2)(e.h, e.h)
    res = (hf.ScfResult * 2)()
    st = (ctypes.c_int * 2)()
    assert hf.lib().hf_scf_batch(hs, 2, P, res, st) == -1
