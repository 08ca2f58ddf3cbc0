"""Independent McMurchie-Davidson ERI / one-electron integrals (golden-vector generator).

This is a second, unrelated algorithm (Hermite expansion + Hermite Coulomb
integrals, Boys from mpmath's 1F1 at 40 digits) used to pin the oracle and the
GPU library.  It shares nothing with the reference's Head-Gordon-Pople scheme
except the definition of the basis functions:

  phi(r) = sum_k c_k N_k (x-Ax)^l (y-Ay)^m (z-Az)^n exp(-a_k |r-A|^2),
  N_k = (2a/pi)^{3/4} (4a)^{L/2} / sqrt((2l-1)!!(2m-1)!!(2n-1)!!)

(reference HartreeFock/GaussianOrbital.cpp:25-30; contractions are not renormalised).
Pure Python/numpy: meant for small molecules only.
"""
import math
from functools import lru_cache

import mpmath
import numpy as np

mpmath.mp.dps = 40


@lru_cache(maxsize=None)
def boys_mp(m, T):
    """F_m(T) = 1F1(m+1/2; m+3/2; -T) / (2m+1), evaluated at 40 digits."""
    T = mpmath.mpf(T)
    return float(mpmath.hyp1f1(m + 0.5, m + 1.5, -T) / (2 * m + 1))


def dfact(n):
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


def norm(a, lmn):
    l, m, n = lmn
    L = l + m + n
    return (2 * a / math.pi) ** 0.75 * (4 * a) ** (L / 2) / math.sqrt(dfact(2 * l - 1) * dfact(2 * m - 1) * dfact(2 * n - 1))


def hermite_E(i, j, t, Qx, a, b):
    """Hermite expansion coefficient E^{ij}_t for one Cartesian direction."""
    p = a + b
    q = a * b / p
    if t < 0 or t > i + j:
        return 0.0
    if i == j == t == 0:
        return math.exp(-q * Qx * Qx)
    if j == 0:
        return (1 / (2 * p)) * hermite_E(i - 1, j, t - 1, Qx, a, b) - (q * Qx / a) * hermite_E(i - 1, j, t, Qx, a, b) + \
            (t + 1) * hermite_E(i - 1, j, t + 1, Qx, a, b)
    return (1 / (2 * p)) * hermite_E(i, j - 1, t - 1, Qx, a, b) + (q * Qx / b) * hermite_E(i, j - 1, t, Qx, a, b) + \
        (t + 1) * hermite_E(i, j - 1, t + 1, Qx, a, b)


def hermite_R_table(tmax, p, PC):
    """R_{tuv}^{(0)} for t+u+v <= tmax by the standard downward-n recursion."""
    X, Y, Z = PC
    T = p * (X * X + Y * Y + Z * Z)
    Rn = {}
    for n in range(tmax + 1):
        Rn[(0, 0, 0, n)] = (-2 * p) ** n * boys_mp(n, T)

    def R(t, u, v, n):
        key = (t, u, v, n)
        if key in Rn:
            return Rn[key]
        if t < 0 or u < 0 or v < 0:
            return 0.0
        if t > 0:
            val = (t - 1) * R(t - 2, u, v, n + 1) + X * R(t - 1, u, v, n + 1)
        elif u > 0:
            val = (u - 1) * R(t, u - 2, v, n + 1) + Y * R(t, u - 1, v, n + 1)
        else:
            val = (v - 1) * R(t, u, v - 2, n + 1) + Z * R(t, u, v - 1, n + 1)
        Rn[key] = val
        return val

    return R


def prim_eri(a, la, A, b, lb, B, c, lc, C, d, ld, D):
    p, q = a + b, c + d
    alpha = p * q / (p + q)
    P = (a * np.asarray(A) + b * np.asarray(B)) / p
    Q = (c * np.asarray(C) + d * np.asarray(D)) / q
    AB = np.asarray(A) - np.asarray(B)
    CD = np.asarray(C) - np.asarray(D)
    Lsum = sum(la) + sum(lb) + sum(lc) + sum(ld)
    R = hermite_R_table(Lsum, alpha, P - Q)
    Eab = [[hermite_E(la[x], lb[x], t, AB[x], a, b) for t in range(la[x] + lb[x] + 1)] for x in range(3)]
    Ecd = [[hermite_E(lc[x], ld[x], t, CD[x], c, d) for t in range(lc[x] + ld[x] + 1)] for x in range(3)]
    val = 0.0
    for t in range(len(Eab[0])):
        for u in range(len(Eab[1])):
            for v in range(len(Eab[2])):
                e1 = Eab[0][t] * Eab[1][u] * Eab[2][v]
                if e1 == 0.0:
                    continue
                for tau in range(len(Ecd[0])):
                    for nu in range(len(Ecd[1])):
                        for phi in range(len(Ecd[2])):
                            e2 = Ecd[0][tau] * Ecd[1][nu] * Ecd[2][phi]
                            sgn = -1.0 if (tau + nu + phi) % 2 else 1.0
                            val += e1 * sgn * e2 * R(t + tau, u + nu, v + phi, 0)
    return 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * val


def prim_overlap_kinetic_nuclear(a, la, A, b, lb, B, charges):
    p = a + b
    P = (a * np.asarray(A) + b * np.asarray(B)) / p
    AB = np.asarray(A) - np.asarray(B)

    def S1(i, j, x):
        return hermite_E(i, j, 0, AB[x], a, b) * math.sqrt(math.pi / p)

    s = [S1(la[x], lb[x], x) for x in range(3)]

    def T1(i, j, x):
        """1-D kinetic term: b(2j+1) S_ij - 2b^2 S_i,j+2 - j(j-1)/2 S_i,j-2."""
        val = b * (2 * j + 1) * S1(i, j, x) - 2 * b * b * S1(i, j + 2, x)
        if j >= 2:
            val -= 0.5 * j * (j - 1) * S1(i, j - 2, x)
        return val

    S = s[0] * s[1] * s[2]
    T = T1(la[0], lb[0], 0) * s[1] * s[2] + s[0] * T1(la[1], lb[1], 1) * s[2] + s[0] * s[1] * T1(la[2], lb[2], 2)
    V = 0.0
    L = sum(la) + sum(lb)
    for Zc, Cpos in charges:
        R = hermite_R_table(L, p, P - np.asarray(Cpos))
        acc = 0.0
        for t in range(la[0] + lb[0] + 1):
            for u in range(la[1] + lb[1] + 1):
                for v in range(la[2] + lb[2] + 1):
                    acc += hermite_E(la[0], lb[0], t, AB[0], a, b) * hermite_E(la[1], lb[1], u, AB[1], a, b) * \
                        hermite_E(la[2], lb[2], v, AB[2], a, b) * R(t, u, v, 0)
        V -= Zc * 2 * math.pi / p * acc
    return S, T, V


class BasisFunction:
    def __init__(self, center, lmn, exps, coefs, atom):
        self.center = tuple(center)
        self.lmn = tuple(lmn)
        self.exps = list(exps)
        self.coefs = list(coefs)
        self.norms = [norm(a, lmn) for a in exps]
        self.atom = atom


def parse_basis(path):
    """NWChem basis reader (same data semantics as the reference: SP shells hold
    s then p sharing exponents; Cartesian components in xx,xy,xz,yy,yz,zz order)."""
    import re
    shell_re = re.compile(r"^(\w+)\s+(\w+)\s*$")
    num_re = re.compile(r"[+-]?[0-9]*\.?[0-9]+((E|D)[+-][0-9]+)?")
    ignore_re = re.compile(r"^\s*(#|BASIS|END).*$|^\s*$")
    symbols = ["H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne"]
    atoms = {}
    cur = None
    for line in open(path):
        line = line.rstrip("\n")
        if ignore_re.match(line):
            continue
        m = shell_re.match(line)
        if m:
            z = symbols.index(m.group(1)) + 1
            cur = {"type": m.group(2).upper(), "exps": [], "coefs": []}
            atoms.setdefault(z, []).append(cur)
            continue
        vals = [float(x.group(0).replace("D", "E")) for x in num_re.finditer(line)]
        cur["exps"].append(vals[0])
        cur["coefs"].append(vals[1:])
    return atoms


def components(L):
    out = []
    for l in range(L, -1, -1):
        for m in range(L - l, -1, -1):
            out.append((l, m, L - l - m))
    return out


def build_basis(path, atoms):
    shells = parse_basis(path)
    bfs = []
    for ai, (z, pos) in enumerate(atoms):
        for sh in shells[z]:
            for li, letter in enumerate(sh["type"]):
                L = "SPDFGH".index(letter)
                coefs = [c[li] for c in sh["coefs"]]
                for lmn in components(L):
                    bfs.append(BasisFunction(pos, lmn, sh["exps"], coefs, ai))
    return bfs


def contracted_eri(f1, f2, f3, f4):
    v = 0.0
    for a, ca, na in zip(f1.exps, f1.coefs, f1.norms):
        for b, cb, nb in zip(f2.exps, f2.coefs, f2.norms):
            for c, cc, nc in zip(f3.exps, f3.coefs, f3.norms):
                for d, cd, nd in zip(f4.exps, f4.coefs, f4.norms):
                    v += ca * cb * cc * cd * na * nb * nc * nd * prim_eri(
                        a, f1.lmn, f1.center, b, f2.lmn, f2.center, c, f3.lmn, f3.center, d, f4.lmn, f4.center)
    return v


def one_electron(bfs, atoms):
    n = len(bfs)
    S, T, V = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    charges = [(z, pos) for z, pos in atoms]
    for i in range(n):
        for j in range(i, n):
            s = t = v = 0.0
            fi, fj = bfs[i], bfs[j]
            for a, ca, na in zip(fi.exps, fi.coefs, fi.norms):
                for b, cb, nb in zip(fj.exps, fj.coefs, fj.norms):
                    ps, pt, pv = prim_overlap_kinetic_nuclear(a, fi.lmn, fi.center, b, fj.lmn, fj.center, charges)
                    w = ca * cb * na * nb
                    s += w * ps
                    t += w * pt
                    v += w * pv
            S[i, j] = S[j, i] = s
            T[i, j] = T[j, i] = t
            V[i, j] = V[j, i] = v
    return S, T, V
