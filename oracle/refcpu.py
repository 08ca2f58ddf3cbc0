"""ctypes view of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Used by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg to check
and time the GPU library; the product never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class ScfParams(ctypes.Structure):
    _fields_ = [("unrestricted", ctypes.c_int), ("max_iterations", ctypes.c_int),
                ("alpha", ctypes.c_double), ("init_guess", ctypes.c_double),
                ("use_diis", ctypes.c_int), ("max_diis_iterations", ctypes.c_int),
                ("normal_iter_after_diis", ctypes.c_int), ("add_asymmetry", ctypes.c_int),
                ("asymmetry", ctypes.c_double)]


class ScfResult(ctypes.Structure):
    _fields_ = [("energy", ctypes.c_double), ("homo", ctypes.c_double),
                ("iterations", ctypes.c_int), ("converged", ctypes.c_int),
                ("s2", ctypes.c_double)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.ref_mol_create.argtypes = [ctypes.c_char_p, ctypes.c_int, ip, dp, ctypes.POINTER(ctypes.c_void_p)]
        L.ref_mol_destroy.argtypes = [ctypes.c_void_p]
        L.ref_mol_nbf.argtypes = [ctypes.c_void_p]
        L.ref_mol_max_l.argtypes = [ctypes.c_void_p]
        L.ref_mol_electrons.argtypes = [ctypes.c_void_p, ip, ip]
        L.ref_mol_nuclear_repulsion.argtypes = [ctypes.c_void_p]
        L.ref_mol_nuclear_repulsion.restype = ctypes.c_double
        L.ref_mol_set_electrons.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.ref_mol_set_field.argtypes = [ctypes.c_void_p, dp]
        L.ref_mol_bf_info.argtypes = [ctypes.c_void_p, ip, ip, ip, ip]
        L.ref_set_exact_transfer.argtypes = [ctypes.c_int]
        L.ref_boys.argtypes = [ctypes.c_int, ctypes.c_double, dp]
        L.ref_eri.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ref_eri.restype = ctypes.c_double
        L.ref_packed_index.argtypes = [ctypes.c_uint32] * 4
        L.ref_packed_index.restype = ctypes.c_uint64
        L.ref_packed_len.argtypes = [ctypes.c_int]
        L.ref_packed_len.restype = ctypes.c_uint64
        L.ref_eri_packed.argtypes = [ctypes.c_void_p, dp, ctypes.c_uint64]
        L.ref_eri_packed_rows.argtypes = [ctypes.c_void_p, dp, ctypes.c_uint64, ctypes.c_int, ctypes.c_int]
        L.ref_eri_packed_sample.argtypes = [ctypes.c_void_p, dp, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                            ctypes.POINTER(ctypes.c_uint64)]
        L.ref_rhf_g.argtypes = [ctypes.c_int, dp, dp, dp, ctypes.c_int, ctypes.c_int]
        L.ref_one_electron.argtypes = [ctypes.c_void_p, dp, dp, dp]
        L.ref_moment.argtypes = [ctypes.c_void_p, dp, dp, dp]
        L.ref_jk_from_packed.argtypes = [ctypes.c_int, dp, dp, dp, dp]
        L.ref_scf_defaults.argtypes = [ctypes.POINTER(ScfParams)]
        L.ref_scf.argtypes = [ctypes.c_void_p, ctypes.POINTER(ScfParams), ctypes.POINTER(ScfResult), dp]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class Molecule:
    """A molecule in the oracle's own data model."""

    def __init__(self, basis_file, atoms):
        z = np.ascontiguousarray([a[0] for a in atoms], dtype=np.int32)
        xyz = np.ascontiguousarray([c for a in atoms for c in a[1]], dtype=np.float64)
        h = ctypes.c_void_p()
        rc = lib().ref_mol_create(basis_file.encode(), len(atoms), _ip(z), _dp(xyz), ctypes.byref(h))
        if rc != 0:
            raise RuntimeError("ref_mol_create failed (%d) for %s" % (rc, basis_file))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().ref_mol_destroy(self.h)
            self.h = None

    @property
    def nbf(self):
        return lib().ref_mol_nbf(self.h)

    def electrons(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        lib().ref_mol_electrons(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def nuclear_repulsion(self):
        return lib().ref_mol_nuclear_repulsion(self.h)

    def set_electrons(self, alpha, beta):
        """Molecule::alphaElectrons / betaElectrons (ions, high-spin states)."""
        if lib().ref_mol_set_electrons(self.h, int(alpha), int(beta)) != 0:
            raise ValueError("bad electron counts")
        return self

    def set_field(self, E):
        """Molecule::ElectricField (au): the SCF's h gains E . M."""
        e = np.ascontiguousarray(E, dtype=np.float64)
        lib().ref_mol_set_field(self.h, _dp(e))
        return self

    def bf_info(self):
        n = self.nbf
        L = np.zeros(n, np.int32)
        lmn = np.zeros(3 * n, np.int32)
        at = np.zeros(n, np.int32)
        sh = np.zeros(n, np.int32)
        lib().ref_mol_bf_info(self.h, _ip(L), _ip(lmn), _ip(at), _ip(sh))
        return L, lmn.reshape(n, 3), at, sh

    def eri(self, i, j, k, l):
        return lib().ref_eri(self.h, i, j, k, l)

    def eri_packed(self, rows=None):
        n = self.nbf
        ln = lib().ref_packed_len(n)
        out = np.zeros(ln, np.float64)
        if rows is None:
            rc = lib().ref_eri_packed(self.h, _dp(out), ln)
        else:
            rc = lib().ref_eri_packed_rows(self.h, _dp(out), ln, rows[0], rows[1])
        if rc != 0:
            raise RuntimeError("ref_eri_packed failed")
        return out

    def eri_packed_sample(self, stride, offset=0, out=None):
        """Every stride-th (i,j,k) triple of the packed loop; returns (array, ERIs evaluated)."""
        ln = lib().ref_packed_len(self.nbf)
        if out is None:
            out = np.zeros(ln, np.float64)
        cnt = ctypes.c_uint64()
        rc = lib().ref_eri_packed_sample(self.h, _dp(out), ln, stride, offset, ctypes.byref(cnt))
        if rc != 0:
            raise RuntimeError("ref_eri_packed_sample failed")
        return out, cnt.value

    def one_electron(self):
        n = self.nbf
        S, T, V = (np.zeros((n, n)) for _ in range(3))
        lib().ref_one_electron(self.h, _dp(S), _dp(T), _dp(V))
        return S, T, V

    def moment(self):
        """Dipole moment integrals about the origin, (Mx, My, Mz) (HartreeFock/QuantumMatrix.cpp:61-90)."""
        n = self.nbf
        M = [np.zeros((n, n)) for _ in range(3)]
        lib().ref_moment(self.h, _dp(M[0]), _dp(M[1]), _dp(M[2]))
        return M

    def scf(self, unrestricted=False, **kw):
        p = ScfParams()
        lib().ref_scf_defaults(ctypes.byref(p))
        p.unrestricted = 1 if unrestricted else 0
        for k, v in kw.items():
            setattr(p, k, v)
        r = ScfResult()
        n = self.nbf
        D = np.zeros((2 if unrestricted else 1, n, n))
        lib().ref_scf(self.h, ctypes.byref(p), ctypes.byref(r), _dp(D))
        return r, D


def set_exact_transfer(on):
    """Switch the oracle between the reference's transfer range (False) and exact ERIs."""
    lib().ref_set_exact_transfer(1 if on else 0)


def boys(maxm, T):
    out = np.zeros(maxm + 1)
    lib().ref_boys(maxm, T, _dp(out))
    return out


def packed_index(i, j, k, l):
    return lib().ref_packed_index(i, j, k, l)


def packed_len(n):
    return lib().ref_packed_len(n)


def rhf_g(n, eri, D, rows=None):
    G = np.zeros((n, n))
    D = np.ascontiguousarray(D, dtype=np.float64)
    i0, i1 = rows if rows is not None else (0, n)
    lib().ref_rhf_g(n, _dp(eri), _dp(D), _dp(G), i0, i1)
    return G


def jk_from_packed(n, eri, D):
    J = np.zeros((n, n))
    K = np.zeros((n, n))
    D = np.ascontiguousarray(D, dtype=np.float64)
    lib().ref_jk_from_packed(n, _dp(eri), _dp(D), _dp(J), _dp(K))
    return J, K


# ---- properties and MP2 (the steps after the SCF) ----

def unpack_eri(n, eri):
    """Full n^4 tensor g[i,j,k,l] from the 8-fold packed store (small n only)."""
    i = np.arange(n)

    def two(a, b):
        hi, lo = np.maximum(a, b), np.minimum(a, b)
        return hi * (hi + 1) // 2 + lo

    ij = two(i[:, None], i[None, :])
    return eri[two(ij[:, :, None, None], ij[None, None, :, :])]


def mulliken_charges(D, S, Z, atom_of_bf):
    """Z_A - sum over the functions of atom A of (D S)_mm, D the total density
    (RestrictedHartreeFock::CalculateAtomicCharge, HartreeFock/RestrictedHartreeFock.cpp:291-317;
    UnrestrictedHartreeFock.cpp:381-406 with D = D_alpha + D_beta)."""
    ds = np.einsum("mn,nm->m", D, S)
    return np.array([z - ds[atom_of_bf == a].sum() for a, z in enumerate(Z)])


def dipole(D, M, Z, R):
    """sum_A Z_A R_A - sum_mn D_mn M_mn (GetMoment, HartreeFock/RestrictedHartreeFock.cpp:319-335,
    GetNuclearMoment, HartreeFock/HartreeFockAlgorithm.cpp:234-242)."""
    nuc = np.einsum("a,ax->x", np.asarray(Z, float), np.asarray(R, float))
    return nuc - np.array([np.sum(D * m) for m in M])


def _orbitals(F, S):
    import scipy.linalg
    return scipy.linalg.eigh(F, S)


def mp2_rhf(n, eri, D, h, S, nocc):
    """RHF MP2 correlation energy as RestrictedHartreeFock::CalculateMp2Energy
    (HartreeFock/RestrictedHartreeFock.cpp:251-288): sum over occupied i, j and virtual a, b of
    (ia|jb) (2 (ia|jb) - (ib|ja)) / (e_i + e_j - e_a - e_b), the orbitals and energies those of the
    converged Fock matrix h + J[D] - K[D]/2."""
    J, K = jk_from_packed(n, eri, D)
    eps, C = _orbitals(h + J - 0.5 * K, S)
    g = unpack_eri(n, eri)
    Co, Cv = C[:, :nocc], C[:, nocc:]
    ovov = np.einsum("pqrs,pi,qa,rj,sb->iajb", g, Co, Cv, Co, Cv, optimize=True)
    eo, ev = eps[:nocc], eps[nocc:]
    den = eo[:, None, None, None] - ev[None, :, None, None] + eo[None, None, :, None] - ev[None, None, None, :]
    return float(np.sum(ovov * (2.0 * ovov - ovov.transpose(0, 3, 2, 1)) / den))


def mp2_uhf_reference(n, eri, Da, Db, h, S, na, nb):
    """UHF MP2 as UnrestrictedHartreeFock::CalculateMp2Energy (HartreeFock/UnrestrictedHartreeFock.cpp:
    295-375): for each spin, sum over same-spin occupied i, j and virtual a, b of
    (ia|jb)^2 / (e_i + e_j - e_a - e_b), the two sums added and multiplied by 1/4.  (This is the
    reference's formula as written: no antisymmetrisation and no opposite-spin term.)"""
    J = jk_from_packed(n, eri, Da)[0] + jk_from_packed(n, eri, Db)[0]
    g = unpack_eri(n, eri)
    tot = 0.0
    for Dsp, no in ((Da, na), (Db, nb)):
        eps, C = _orbitals(h + J - jk_from_packed(n, eri, Dsp)[1], S)
        Co, Cv = C[:, :no], C[:, no:]
        ovov = np.einsum("pqrs,pi,qa,rj,sb->iajb", g, Co, Cv, Co, Cv, optimize=True)
        eo, ev = eps[:no], eps[no:]
        den = eo[:, None, None, None] - ev[None, :, None, None] + eo[None, None, :, None] - ev[None, None, None, :]
        tot += float(np.sum(ovov * ovov / den))
    return 0.25 * tot
