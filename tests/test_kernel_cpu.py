"""CPU check of the generated ERI class code (the same eval_quartet() the GPU kernels run,
compiled for the host by the test harness tests/cpu_kernel/quartet_host.hip) against the
oracle, for every angular class, on molecules whose pairs are two-centre."""
import ctypes
import math
import os

import numpy as np
import pytest

from conftest import basis
from hartreefock_amd import geometry as g
from oracle import refcpu
import md_eri

HERE = os.path.dirname(os.path.abspath(__file__))
_H = None


def harness():
    global _H
    if _H is None:
        path = os.path.join(HERE, "cpu_kernel", "libquartet_host.so")
        if not os.path.exists(path):
            pytest.skip("host harness not built (python -m hartreefock_amd.build)")
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        L.hfq_eval.argtypes = [ctypes.c_int] * 5 + [dp] * 4 + [dp, ctypes.c_int, dp, ctypes.c_int, dp, dp] + \
            [ctypes.c_int] * 6
        L.hfq_boys_table.argtypes = [dp]
        fp = ctypes.POINTER(ctypes.c_float)
        L.hfq_eval_f32.argtypes = [ctypes.c_int] * 4 + [dp] * 4 + [dp, ctypes.c_int, dp, ctypes.c_int, fp, fp]
        L.hfq_boys.argtypes = [ctypes.c_int, ctypes.c_double, dp, dp]
        L.hfq_boys_f32.argtypes = [ctypes.c_int, ctypes.c_float, fp, fp]
        _H = L
    return _H


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def shells_of(bfile, atoms):
    """Shells after splitting SP, in the reference's basis-function order."""
    sh = md_eri.parse_basis(bfile)
    out = []
    bf = 0
    for ai, (z, pos) in enumerate(atoms):
        for s in sh[z]:
            for li, letter in enumerate(s["type"]):
                L = "SPD".index(letter)
                coef = [c[li] * (2 * a / math.pi) ** 0.75 * (4 * a) ** (L / 2.0) for a, c in zip(s["exps"], s["coefs"])]
                out.append(dict(L=L, xyz=np.array(pos, float), alpha=list(s["exps"]), coef=coef, bf=bf, atom=ai,
                                index=len(out)))
                bf += (L + 1) * (L + 2) // 2
    return out


def prim_pairs(A, B):
    """64-byte PrimPair records, as the library builds them (csrc/context.hip)."""
    rec = []
    R2 = float(((A["xyz"] - B["xyz"]) ** 2).sum())
    for a, ca in zip(A["alpha"], A["coef"]):
        for b, cb in zip(B["alpha"], B["coef"]):
            z = a + b
            P = (a * A["xyz"] + b * B["xyz"]) / z
            K = ca * cb * math.exp(-a * b / z * R2)
            rec.append([P[0], P[1], P[2], z, K / z, 1.0 / z, 0.0, 0.0])
    return np.ascontiguousarray(rec, dtype=np.float64)


def block(A, B, C, D, exact, boys):
    nA, nB, nC, nD = (((s["L"] + 1) * (s["L"] + 2)) // 2 for s in (A, B, C, D))
    out = np.zeros(nA * nB * nC * nD)
    bp, kp = prim_pairs(A, B), prim_pairs(C, D)
    rc = harness().hfq_eval(A["L"], B["L"], C["L"], D["L"], exact, _dp(A["xyz"]), _dp(B["xyz"]), _dp(C["xyz"]),
                            _dp(D["xyz"]), _dp(bp), len(bp), _dp(kp), len(kp), _dp(boys), _dp(out),
                            A["index"], C["index"], A["bf"], B["bf"], C["bf"], D["bf"])
    assert rc == 0
    return out.reshape(nA, nB, nC, nD)


def boys_table():
    t = np.zeros(harness().hfq_boys_table_size())
    harness().hfq_boys_table(_dp(t))
    return t


def _boys_golden():
    rows = [l.split() for l in open(os.path.join(HERE, "golden", "boys_mpmath.txt")) if not l.startswith("#")]
    table = {}
    for T, m, v in rows:
        table.setdefault(float(T), {})[int(m)] = float(v)
    return table


def test_kernel_boys_matches_mpmath():
    """The kernels' Boys evaluation (table of F_m(0.1 k) and e^{-0.1 k}, 8-term Taylor expansion of F_L about the
    nearest grid point and downward recursion below T = 36; asymptotic F_0 and upward recursion above), for every
    top order L <= 8 a quartet class uses, against 40-digit mpmath at every committed T (0 to 1000, including
    35.95, 36 and 36.07 on both sides of the switch): FP64 within 2e-14 relative, the mixed tier's FP32
    evaluation within 2e-6."""
    tab = boys_table()
    tab32 = tab.astype(np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    worst, worst32 = 0.0, 0.0
    for T, vals in _boys_golden().items():
        for L in range(9):
            F = np.zeros(L + 1)
            assert harness().hfq_boys(L, T, _dp(tab), _dp(F)) == 0
            F32 = np.zeros(L + 1, np.float32)
            assert harness().hfq_boys_f32(L, T, tab32.ctypes.data_as(fp), F32.ctypes.data_as(fp)) == 0
            for m in range(L + 1):
                worst = max(worst, abs(F[m] - vals[m]) / vals[m])
                worst32 = max(worst32, abs(float(F32[m]) - vals[m]) / vals[m])
    print("Boys FP64 worst relative %.2e, FP32 %.2e" % (worst, worst32))
    assert worst < 2e-14, worst
    assert worst32 < 2e-6, worst32


CASES = [("ethane_631gs", "6-31g_st_.1.nw", g.alkane_zigzag(2, seed=5)),
         ("c2_631gs", "6-31g_st_.1.nw", [(6, (0.0, 0.0, 0.0)), (6, (0.3, 1.2, 2.5))])]


@pytest.mark.parametrize("name,bfile,atoms", CASES, ids=[c[0] for c in CASES])
def test_every_class_matches_oracle_exact(name, bfile, atoms):
    """Exact convention, all canonical classes (La>=Lb, Lc>=Ld, bra L-sum >= ket L-sum), a seeded
    sample of shell quartets per class, every component against the oracle (exact mode)."""
    shells = shells_of(basis(bfile), atoms)
    mol = refcpu.Molecule(basis(bfile), atoms)
    boys = boys_table()
    rng = np.random.default_rng(3)
    by_l = {}
    for i, s in enumerate(shells):
        by_l.setdefault(s["L"], []).append(i)
    pairs = [(a, b) for a in range(3) for b in range(a + 1)]
    worst = 0.0
    refcpu.set_exact_transfer(True)
    try:
        for (la, lb) in pairs:
            for (lc, ld) in pairs:
                if (lc + ld, lc) > (la + lb, la):
                    continue
                if not (by_l.get(la) and by_l.get(lb) and by_l.get(lc) and by_l.get(ld)):
                    continue
                for _ in range(3):
                    A, B, C, D = (shells[int(rng.choice(by_l[x]))] for x in (la, lb, lc, ld))
                    blk = block(A, B, C, D, 1, boys)
                    for ia in range(blk.shape[0]):
                        for ib in range(blk.shape[1]):
                            for ic in range(blk.shape[2]):
                                for id_ in range(blk.shape[3]):
                                    ref = mol.eri(A["bf"] + ia, B["bf"] + ib, C["bf"] + ic, D["bf"] + id_)
                                    worst = max(worst, abs(blk[ia, ib, ic, id_] - ref))
    finally:
        refcpu.set_exact_transfer(False)
    assert worst < 1e-12, worst


@pytest.mark.parametrize("name,bfile,atoms", CASES, ids=[c[0] for c in CASES])
def test_dddd_reference_convention_matches_oracle(name, bfile, atoms):
    """Reference convention for (dd|dd): every shell quartet of d shells (pairs stored with the
    higher shell index first, as the library does), every component, against the oracle in its
    default (reference) mode, which canonicalises each component quartet as the reference does."""
    shells = shells_of(basis(bfile), atoms)
    mol = refcpu.Molecule(basis(bfile), atoms)
    boys = boys_table()
    d = [s for s in shells if s["L"] == 2]
    pairs = [(x, y) for x in d for y in d if x["index"] >= y["index"]]
    worst = 0.0
    for pi, (A, B) in enumerate(pairs):
        for (C, D) in pairs[: pi + 1]:
            for (P, Q) in (((A, B), (C, D)), ((C, D), (A, B))):    # both launch orientations
                blk = block(P[0], P[1], Q[0], Q[1], 0, boys)
                for ia in range(6):
                    for ib in range(6):
                        for ic in range(6):
                            for id_ in range(6):
                                ref = mol.eri(P[0]["bf"] + ia, P[1]["bf"] + ib, Q[0]["bf"] + ic, Q[1]["bf"] + id_)
                                worst = max(worst, abs(blk[ia, ib, ic, id_] - ref))
    assert worst < 1e-12, worst


def _coop_eval(cg, A, B, C, D, boys):
    """Interpret a cooperative class's op tables (what the wavefront kernel executes) in numpy."""
    import sys
    csrc = os.path.join(os.path.dirname(HERE), "hartreefock_amd", "csrc")
    if csrc not in sys.path:
        sys.path.insert(0, csrc)
    val = np.zeros(cg.nslot)
    L = cg.g.L
    for P in prim_pairs(A, B):
        for Q in prim_pairs(C, D):
            inv_s = 1.0 / (P[3] + Q[3])
            rho = P[3] * Q[3] * inv_s
            PQ = P[:3] - Q[:3]
            T = rho * float(PQ @ PQ)
            pref = 2 * math.pi ** 2.5 * P[4] * Q[4] * math.sqrt(inv_s)
            F = [pref * md_eri.boys_mp(m, T) for m in range(L + 1)]
            for m in range(L + 1):
                val[cg.F0 + m] = F[m]
            PA = P[:3] - A["xyz"]
            rz = Q[3] * inv_s
            zoe = P[3] * Q[5]
            cf = np.zeros(48)
            cf[1] = 1.0
            cf[2:5] = PA
            cf[5:8] = -rz * PQ
            cf[8:11] = zoe * PA + (Q[:3] - C["xyz"])
            cf[11] = -zoe
            for k in range(1, 9):
                cf[11 + k] = k * 0.5 * P[5]
                cf[19 + k] = -rz * k * 0.5 * P[5]
                cf[27 + k] = k * 0.5 * Q[5]
            for dst, src, cs in cg.p_ops:
                val[dst] = sum(cf[c] * val[s] for s, c in zip(src, cs))
    cf = np.zeros(48)
    cf[1] = 1.0
    cf[36:39] = A["xyz"] - B["xyz"]
    cf[39:42] = C["xyz"] - D["xyz"]
    for dst, src, cs in cg.h_ops:
        val[dst] = sum(cf[c] * val[s] for s, c in zip(src, cs))
    return np.array([val[s] * nv for s, nv in zip(cg.out_slot, cg.out_norm)])


def test_cooperative_op_tables_match_oracle():
    """The op lists the one-wavefront-per-quartet kernels execute, interpreted level by level in
    numpy (Boys from mpmath), against the oracle (exact convention) for every cooperative class."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(HERE), "hartreefock_amd", "csrc"))
    import gen_eri
    atoms = [(6, (0.0, 0.0, 0.0)), (6, (0.3, 1.2, 2.5))]
    bfile = basis("6-31g_st_.1.nw")
    shells = shells_of(bfile, atoms)
    mol = refcpu.Molecule(bfile, atoms)
    by = {}
    for s in shells:
        by.setdefault((s["L"], s["atom"]), []).append(s)
    refcpu.set_exact_transfer(True)
    worst = 0.0
    try:
        for cls in [(2, 1, 2, 1), (2, 2, 1, 1), (2, 2, 2, 0), (2, 2, 2, 1), (2, 2, 2, 2)]:
            cg = gen_eri.CoopGen(*cls, True).build()
            A = by[(cls[0], 1)][0]
            B = by[(cls[1], 0)][-1]
            C = by[(cls[2], 0)][0]
            D = by[(cls[3], 1)][-1]
            blk = _coop_eval(cg, A, B, C, D, None)
            nb_, nc_, nd_ = ((x + 1) * (x + 2) // 2 for x in cls[1:])
            for idx, v in enumerate(blk):
                id_ = idx % nd_
                ic = (idx // nd_) % nc_
                ib = (idx // (nd_ * nc_)) % nb_
                ia = idx // (nd_ * nc_ * nb_)
                ref = mol.eri(A["bf"] + ia, B["bf"] + ib, C["bf"] + ic, D["bf"] + id_)
                worst = max(worst, abs(v - ref))
    finally:
        refcpu.set_exact_transfer(False)
    assert worst < 1e-12, worst


def test_cooperative_truncated_dddd_matches_oracle_reference_mode():
    """The reference-convention (dd|dd) op tables with the kernel's per-component bra choice
    (both launch orientations, one- and two-centre pairs) against the oracle's default mode."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(HERE), "hartreefock_amd", "csrc"))
    import gen_eri
    atoms = [(6, (0.0, 0.0, 0.0)), (6, (0.3, 1.2, 2.5))]
    bfile = basis("6-31g_st_.1.nw")
    shells = shells_of(bfile, atoms)
    mol = refcpu.Molecule(bfile, atoms)
    cg = gen_eri.CoopGen(2, 2, 2, 2, False).build()
    d = [s for s in shells if s["L"] == 2]
    pairs = [(x, y) for x in d for y in d if x["index"] >= y["index"]]
    worst = 0.0
    for (A, B) in pairs:
        for (C, D) in pairs:
            needP, needQ = A["index"] >= C["index"], A["index"] <= C["index"]
            res = np.zeros(1296)
            if needP:
                res[:] = _coop_eval(cg, A, B, C, D, None)
            if needQ:
                t = _coop_eval(cg, C, D, A, B, None)
                for idx in range(1296):
                    id_, ic, ib, ia = idx % 6, (idx // 6) % 6, (idx // 36) % 6, idx // 216
                    a, b, c, dd = A["bf"] + ia, B["bf"] + ib, C["bf"] + ic, D["bf"] + id_
                    i, j, k, l = max(a, b), min(a, b), max(c, dd), min(c, dd)
                    braP = needP and not (i < k or (i == k and j < l))
                    if not braP:
                        res[idx] = t[((ic * 6 + id_) * 6 + ia) * 6 + ib]
            for idx in range(0, 1296, 7):
                id_, ic, ib, ia = idx % 6, (idx // 6) % 6, (idx // 36) % 6, idx // 216
                ref = mol.eri(A["bf"] + ia, B["bf"] + ib, C["bf"] + ic, D["bf"] + id_)
                worst = max(worst, abs(res[idx] - ref))
    assert worst < 1e-12, worst


def block_f32(A, B, C, D, boys32):
    nA, nB, nC, nD = (((s["L"] + 1) * (s["L"] + 2)) // 2 for s in (A, B, C, D))
    out = np.zeros(nA * nB * nC * nD, dtype=np.float32)
    bp, kp = prim_pairs(A, B), prim_pairs(C, D)
    fp = ctypes.POINTER(ctypes.c_float)
    rc = harness().hfq_eval_f32(A["L"], B["L"], C["L"], D["L"], _dp(A["xyz"]), _dp(B["xyz"]), _dp(C["xyz"]),
                                _dp(D["xyz"]), _dp(bp), len(bp), _dp(kp), len(kp), boys32.ctypes.data_as(fp),
                                out.ctypes.data_as(fp))
    assert rc == 0
    return out.astype(np.float64).reshape(nA, nB, nC, nD)


def schwarz(A, B, boys):
    """Schwarz factor sqrt(max |(ab|ab)|) of a shell pair (FP64 harness)."""
    d = block(A, B, A, B, 1, boys)
    n1, n2 = d.shape[0], d.shape[1]
    return math.sqrt(max(abs(d[i, j, i, j]) for i in range(n1) for j in range(n2)))


@pytest.mark.parametrize("name,bfile,atoms", CASES, ids=[c[0] for c in CASES])
def test_fp32_classes_relative_accuracy(name, bfile, atoms):
    """The FP32 evaluation used for small quartets in the mixed-precision tier (every class the
    thread-per-quartet kernels serve, L < 6), against the oracle's FP64 integrals.  FP32 rounding
    (6e-8) is amplified by the electron-transfer step, whose (zeta/eta) factor reaches 1e3-1e4 for
    a tight bra primitive against a diffuse ket one (the same cancellation costs FP64 about four
    digits); the error of a shell quartet stays below 1e-3 of its Schwarz bound Q_ab Q_cd, which
    for the quartets under the 1e-6 bound that select FP32 is below 1e-9 absolute.  Classes with
    an s-type ket (no transfer step) stay near FP32 rounding (< 1e-5 of the bound)."""
    shells = shells_of(basis(bfile), atoms)
    mol = refcpu.Molecule(basis(bfile), atoms)
    boys = boys_table()
    boys32 = boys.astype(np.float32)
    rng = np.random.default_rng(11)
    by_l = {}
    for i, s in enumerate(shells):
        by_l.setdefault(s["L"], []).append(i)
    pairs = [(a, b) for a in range(3) for b in range(a + 1)]
    worst = 0.0
    worst_low = 0.0
    checked = 0
    refcpu.set_exact_transfer(True)
    try:
        for (la, lb) in pairs:
            for (lc, ld) in pairs:
                if (lc + ld, lc) > (la + lb, la) or la + lb + lc + ld >= 6:
                    continue
                if not (by_l.get(la) and by_l.get(lb) and by_l.get(lc) and by_l.get(ld)):
                    continue
                for _ in range(3):
                    A, B, C, D = (shells[int(rng.choice(by_l[x]))] for x in (la, lb, lc, ld))
                    blk = block_f32(A, B, C, D, boys32)
                    ref = np.zeros_like(blk)
                    for idx in np.ndindex(*blk.shape):
                        ref[idx] = mol.eri(A["bf"] + idx[0], B["bf"] + idx[1], C["bf"] + idx[2], D["bf"] + idx[3])
                    scale = schwarz(A, B, boys) * schwarz(C, D, boys)
                    err = np.abs(blk - ref).max() / scale
                    worst = max(worst, err)
                    if lc + ld == 0:
                        worst_low = max(worst_low, err)
                    checked += 1
    finally:
        refcpu.set_exact_transfer(False)
    assert checked >= 10
    assert worst_low < 1e-5, worst_low
    assert worst < 1e-3, worst
