"""ctypes binding of libhfgpu.so (include/hfgpu.h)."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

# every function include/hfgpu.h declares
EXPORTED_SYMBOLS = [
    "hf_default_options", "hf_create", "hf_destroy", "hf_last_error", "hf_set_basis", "hf_set_molecule",
    "hf_nbf", "hf_nshell", "hf_electrons", "hf_nuclear_repulsion", "hf_set_electrons", "hf_set_electric_field",
    "hf_nuclear_electric_field_energy", "hf_packed_index", "hf_packed_len",
    "hf_eri_packed", "hf_eri_packed_device", "hf_build_jk", "hf_build_jk_device", "hf_one_electron",
    "hf_pair_matrix_len", "hf_eri_pairs_device", "hf_build_jk_incore_device", "hf_build_jk_incore",
    "hf_build_jk_incore_device_noalloc",
    "hf_properties", "hf_core_hamiltonian", "hf_boys",
    "hf_scf_defaults", "hf_scf", "hf_scf_begin", "hf_scf_step", "hf_scf_end", "hf_scf_batch", "hf_kernel_stats",
    "hf_kernel_stats_reset",
    "hf_serve_default_options", "hf_serve_create", "hf_serve_submit", "hf_serve_poll", "hf_serve_pending",
    "hf_serve_footprint", "hf_serve_destroy", "hf_serve_kernel_stats",
]


class HFError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("hfgpu error %d: %s" % (code, msg))
        self.code = code


class Options(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("schwarz_threshold", ctypes.c_double),
                ("prim_threshold", ctypes.c_double), ("exact_transfer", ctypes.c_int),
                ("precision", ctypes.c_int), ("fp32_threshold", ctypes.c_double), ("profile", ctypes.c_int),
                ("profile_class", ctypes.c_int), ("incore_shape", ctypes.c_int), ("deterministic", ctypes.c_int)]


class BasisSoa(ctypes.Structure):
    _fields_ = [("nshell", ctypes.c_int), ("nbf", ctypes.c_int),
                ("L", ctypes.POINTER(ctypes.c_int)), ("nprim", ctypes.POINTER(ctypes.c_int)),
                ("prim_offset", ctypes.POINTER(ctypes.c_int)), ("bf_offset", ctypes.POINTER(ctypes.c_int)),
                ("atom", ctypes.POINTER(ctypes.c_int)), ("xyz", ctypes.POINTER(ctypes.c_double)),
                ("alpha", ctypes.POINTER(ctypes.c_double)), ("coef", ctypes.POINTER(ctypes.c_double)),
                ("natoms", ctypes.c_int), ("Z", ctypes.POINTER(ctypes.c_int)),
                ("atom_xyz", ctypes.POINTER(ctypes.c_double))]


class ScfParams(ctypes.Structure):
    _fields_ = [("unrestricted", ctypes.c_int), ("max_iterations", ctypes.c_int), ("alpha", ctypes.c_double),
                ("init_guess", ctypes.c_double), ("use_diis", ctypes.c_int), ("max_diis_iterations", ctypes.c_int),
                ("normal_iter_after_diis", ctypes.c_int), ("add_asymmetry", ctypes.c_int),
                ("asymmetry", ctypes.c_double), ("energy_tolerance", ctypes.c_double),
                ("density_tolerance", ctypes.c_double), ("diis_tolerance", ctypes.c_double),
                ("terminate", ctypes.POINTER(ctypes.c_int)), ("density_method", ctypes.c_int),
                ("incremental_fock", ctypes.c_int), ("full_rebuild_every", ctypes.c_int),
                ("final_fp64_iterations", ctypes.c_int), ("incore", ctypes.c_int),
                ("incore_max_bytes", ctypes.c_double), ("mp2", ctypes.c_int)]


class ScfResult(ctypes.Structure):
    _fields_ = [("energy", ctypes.c_double), ("homo", ctypes.c_double), ("s2", ctypes.c_double),
                ("iterations", ctypes.c_int), ("converged", ctypes.c_int), ("jk_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double), ("fp32_builds", ctypes.c_int), ("diag_steps", ctypes.c_int),
                ("full_builds", ctypes.c_int), ("incremental_builds", ctypes.c_int), ("quartets", ctypes.c_double),
                ("sp2_steps", ctypes.c_int), ("incore", ctypes.c_int), ("eri_ms", ctypes.c_double),
                ("mp2_energy", ctypes.c_double)]


class ServeOptions(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_int), ("hbm_budget", ctypes.c_double), ("ctx", Options), ("scf", ScfParams),
                ("lockstep", ctypes.c_int), ("prep_threads", ctypes.c_int)]


class ServeResult(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("tag", ctypes.c_uint64), ("status", ctypes.c_int), ("nbf", ctypes.c_int),
                ("energy", ctypes.c_double), ("iterations", ctypes.c_int), ("converged", ctypes.c_int),
                ("arrival_ms", ctypes.c_double), ("admit_ms", ctypes.c_double), ("done_ms", ctypes.c_double),
                ("footprint", ctypes.c_double)]


class KernelStat(ctypes.Structure):
    _fields_ = [("la", ctypes.c_int), ("lb", ctypes.c_int), ("lc", ctypes.c_int), ("ld", ctypes.c_int),
                ("launches", ctypes.c_int64), ("quartets", ctypes.c_int64), ("prim_quartets", ctypes.c_int64),
                ("eris", ctypes.c_double), ("flops", ctypes.c_double), ("ms", ctypes.c_double),
                ("fp32", ctypes.c_int)]

    def name(self):
        s = "spdfg"
        if self.la < 0:
            return "incore_jk" + ("_f32" if self.fp32 else "")
        return "(%s%s|%s%s)%s" % (s[self.la], s[self.lb], s[self.lc], s[self.ld], "f32" if self.fp32 else "")


def library_path():
    return os.path.join(_HERE, "libhfgpu.so")


def lib():
    """Load libhfgpu.so; raises if it has not been built (no fallback path exists)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not os.path.exists(path):
        raise RuntimeError("libhfgpu.so not built: run `python -m hartreefock_amd.build` (%s)" % path)
    L = ctypes.CDLL(path)
    vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    L.hf_default_options.argtypes = [ctypes.POINTER(Options)]
    L.hf_default_options.restype = None
    L.hf_create.argtypes = [ctypes.POINTER(Options), ctypes.POINTER(vp)]
    L.hf_destroy.argtypes = [vp]
    L.hf_last_error.argtypes = [vp]
    L.hf_last_error.restype = ctypes.c_char_p
    L.hf_set_basis.argtypes = [vp, ctypes.POINTER(BasisSoa)]
    L.hf_set_molecule.argtypes = [vp, ctypes.c_char_p, ctypes.c_int, ip, dp]
    L.hf_nbf.argtypes = [vp]
    L.hf_nshell.argtypes = [vp]
    L.hf_electrons.argtypes = [vp, ip, ip]
    L.hf_nuclear_repulsion.argtypes = [vp]
    L.hf_nuclear_repulsion.restype = ctypes.c_double
    L.hf_set_electrons.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.hf_set_electric_field.argtypes = [vp, dp]
    L.hf_nuclear_electric_field_energy.argtypes = [vp]
    L.hf_nuclear_electric_field_energy.restype = ctypes.c_double
    L.hf_packed_index.argtypes = [ctypes.c_uint32] * 4
    L.hf_packed_index.restype = ctypes.c_uint64
    L.hf_packed_len.argtypes = [ctypes.c_int]
    L.hf_packed_len.restype = ctypes.c_uint64
    L.hf_eri_packed.argtypes = [vp, dp, ctypes.c_uint64]
    L.hf_eri_packed_device.argtypes = [vp, vp, ctypes.c_uint64, vp]
    L.hf_build_jk.argtypes = [vp, ctypes.c_int, dp, dp, dp]
    L.hf_build_jk_device.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp]
    L.hf_pair_matrix_len.argtypes = [ctypes.c_int]
    L.hf_pair_matrix_len.restype = ctypes.c_uint64
    L.hf_eri_pairs_device.argtypes = [vp, vp, ctypes.c_uint64, vp]
    L.hf_build_jk_incore_device.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp, vp]
    L.hf_build_jk_incore.argtypes = [vp, ctypes.c_int, dp, dp, dp]
    L.hf_build_jk_incore_device_noalloc.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp, vp]
    L.hf_one_electron.argtypes = [vp, dp, dp, dp]
    L.hf_core_hamiltonian.argtypes = [vp, dp]
    L.hf_boys.argtypes = [ctypes.c_int, ctypes.c_double, dp]
    L.hf_properties.argtypes = [vp, ctypes.c_int, dp, dp, dp]
    L.hf_scf_defaults.argtypes = [ctypes.POINTER(ScfParams)]
    L.hf_scf_defaults.restype = None
    L.hf_scf.argtypes = [vp, ctypes.POINTER(ScfParams), ctypes.POINTER(ScfResult), dp]
    L.hf_scf_begin.argtypes = [vp, ctypes.POINTER(ScfParams)]
    L.hf_scf_step.argtypes = [vp, ctypes.c_int, dp, dp, dp]
    L.hf_scf_end.argtypes = [vp, ctypes.POINTER(ScfResult), dp]
    L.hf_scf_batch.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(ScfParams),
                               ctypes.POINTER(ScfResult), ip]
    L.hf_kernel_stats.argtypes = [vp, ctypes.POINTER(KernelStat), ctypes.c_int]
    L.hf_kernel_stats_reset.argtypes = [vp]
    L.hf_serve_default_options.argtypes = [ctypes.POINTER(ServeOptions)]
    L.hf_serve_default_options.restype = None
    L.hf_serve_create.argtypes = [ctypes.POINTER(ServeOptions), ctypes.POINTER(vp)]
    L.hf_serve_submit.argtypes = [vp, ctypes.c_char_p, ctypes.c_int, ip, dp, ctypes.c_int, ctypes.c_uint64,
                                  ctypes.POINTER(ctypes.c_uint64)]
    L.hf_serve_poll.argtypes = [vp, ctypes.POINTER(ServeResult), ctypes.c_int, ctypes.c_int]
    L.hf_serve_pending.argtypes = [vp]
    L.hf_serve_kernel_stats.argtypes = [vp, ctypes.POINTER(KernelStat), ctypes.c_int, ctypes.c_int]
    L.hf_serve_footprint.argtypes = [ctypes.c_int, ctypes.c_int]
    L.hf_serve_footprint.restype = ctypes.c_double
    L.hf_serve_destroy.argtypes = [vp]
    _LIB = L
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def packed_index(i, j, k, l):
    return lib().hf_packed_index(i, j, k, l)


def packed_len(n):
    return lib().hf_packed_len(n)


class Engine:
    """One hf_ctx: a molecule on one GPU."""

    def __init__(self, **opts):
        L = lib()
        o = Options()
        L.hf_default_options(ctypes.byref(o))
        for k, v in opts.items():
            if not hasattr(o, k):
                raise TypeError("unknown option %s" % k)
            setattr(o, k, v)
        self.options = o
        h = ctypes.c_void_p()
        rc = L.hf_create(ctypes.byref(o), ctypes.byref(h))
        if rc != 0:
            raise HFError(rc, "hf_create failed (is a GPU visible?)")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().hf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise HFError(rc, lib().hf_last_error(self.h).decode())

    def set_molecule(self, basis_file, atoms):
        self._natoms = len(atoms)
        z = np.ascontiguousarray([a[0] for a in atoms], dtype=np.int32)
        xyz = np.ascontiguousarray([c for a in atoms for c in a[1]], dtype=np.float64)
        self._check(lib().hf_set_molecule(self.h, basis_file.encode(), len(atoms), _ip(z), _dp(xyz)))
        return self

    def set_basis(self, shells, atoms):
        """Install a molecule from shell arrays (hf_set_basis): shells is a list of
        (L, atom index, [exponents], [coefficient x radial normalisation]) with SP already split, in the
        basis-function order; atoms is [(Z, (x, y, z)), ...] in bohr."""
        self._natoms = len(atoms)
        L = np.ascontiguousarray([s[0] for s in shells], dtype=np.int32)
        at = np.ascontiguousarray([s[1] for s in shells], dtype=np.int32)
        npr = np.ascontiguousarray([len(s[2]) for s in shells], dtype=np.int32)
        poff = np.ascontiguousarray(np.concatenate([[0], np.cumsum(npr)[:-1]]), dtype=np.int32)
        nc = (L + 1) * (L + 2) // 2
        bfo = np.ascontiguousarray(np.concatenate([[0], np.cumsum(nc)[:-1]]), dtype=np.int32)
        al = np.ascontiguousarray([x for s in shells for x in s[2]], dtype=np.float64)
        co = np.ascontiguousarray([x for s in shells for x in s[3]], dtype=np.float64)
        z = np.ascontiguousarray([a[0] for a in atoms], dtype=np.int32)
        axyz = np.ascontiguousarray([c for a in atoms for c in a[1]], dtype=np.float64)
        sxyz = np.ascontiguousarray([c for s in shells for c in atoms[s[1]][1]], dtype=np.float64)
        b = BasisSoa(len(shells), int(nc.sum()), _ip(L), _ip(npr), _ip(poff), _ip(bfo), _ip(at), _dp(sxyz), _dp(al),
                     _dp(co), len(atoms), _ip(z), _dp(axyz))
        self._check(lib().hf_set_basis(self.h, ctypes.byref(b)))
        return self

    def set_electrons(self, alpha, beta):
        """Molecule::alphaElectrons / betaElectrons of the installed molecule (0, 0: neutral)."""
        self._check(lib().hf_set_electrons(self.h, int(alpha), int(beta)))
        return self

    def set_electric_field(self, field):
        """Molecule::ElectricField (au) of the installed molecule."""
        f = np.ascontiguousarray(field, dtype=np.float64)
        self._check(lib().hf_set_electric_field(self.h, _dp(f)))
        return self

    def nuclear_electric_field_energy(self):
        return lib().hf_nuclear_electric_field_energy(self.h)

    @property
    def nbf(self):
        return lib().hf_nbf(self.h)

    @property
    def nshell(self):
        return lib().hf_nshell(self.h)

    def electrons(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        lib().hf_electrons(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def nuclear_repulsion(self):
        return lib().hf_nuclear_repulsion(self.h)

    def eri_packed(self):
        n = self.nbf
        ln = lib().hf_packed_len(n)
        out = np.zeros(ln, np.float64)
        self._check(lib().hf_eri_packed(self.h, _dp(out), ln))
        return out

    def build_jk(self, D):
        D = np.ascontiguousarray(D, dtype=np.float64)
        n = self.nbf
        nd = 1 if D.ndim == 2 else D.shape[0]
        D3 = D.reshape(nd, n, n)
        J = np.zeros((n, n))
        K = np.zeros((nd, n, n))
        self._check(lib().hf_build_jk(self.h, nd, _dp(D3), _dp(J), _dp(K)))
        return J, (K[0] if D.ndim == 2 else K)

    def build_jk_device(self, d_D, d_J, d_K, ndens=1, stream=None):
        """Device pointers (ints) in, device pointers out; stream is a hipStream_t as int (None: the engine's)."""
        self._check(lib().hf_build_jk_device(self.h, ndens, ctypes.c_void_p(d_D), ctypes.c_void_p(d_J),
                                             ctypes.c_void_p(d_K), ctypes.c_void_p(stream or 0)))

    def build_jk_incore(self, D):
        """J/K from the in-core ERI store (evaluated in the engine's device memory first)."""
        D = np.ascontiguousarray(D, dtype=np.float64)
        n = self.nbf
        nd = 1 if D.ndim == 2 else D.shape[0]
        D3 = D.reshape(nd, n, n)
        J = np.zeros((n, n))
        K = np.zeros((nd, n, n))
        self._check(lib().hf_build_jk_incore(self.h, nd, _dp(D3), _dp(J), _dp(K)))
        return J, (K[0] if D.ndim == 2 else K)

    def build_jk_incore_device(self, d_G, d_D, d_J, d_K, ndens=1, stream=None, no_alloc=False):
        """In-core J/K from the device store d_G (hf_eri_pairs_device layout); device pointers as ints.  no_alloc:
        the allocation-free entry point (n <= 512, scratch shared by every call on the GPU)."""
        f = lib().hf_build_jk_incore_device_noalloc if no_alloc else lib().hf_build_jk_incore_device
        self._check(f(self.h, ndens, ctypes.c_void_p(d_G), ctypes.c_void_p(d_D), ctypes.c_void_p(d_J),
                      ctypes.c_void_p(d_K), ctypes.c_void_p(stream or 0)))

    def eri_packed_device(self, d_out, length, stream=None):
        """Packed unique ERIs into the device buffer d_out (hf_packed_len(n) doubles)."""
        self._check(lib().hf_eri_packed_device(self.h, ctypes.c_void_p(d_out), length, ctypes.c_void_p(stream or 0)))

    def eri_pairs_device(self, d_G, length, stream=None):
        """Fill the device pair matrix G (hf_pair_matrix_len(n) doubles) with the ERIs."""
        self._check(lib().hf_eri_pairs_device(self.h, ctypes.c_void_p(d_G), length, ctypes.c_void_p(stream or 0)))

    def properties(self, D):
        """Mulliken charges (per atom) and dipole moment (au) of D: the RHF total density (n x n) or the
        UHF alpha and beta densities (2 x n x n)."""
        D = np.ascontiguousarray(D, dtype=np.float64)
        nd = 1 if D.ndim == 2 else D.shape[0]
        q = np.zeros(self._natoms)
        mu = np.zeros(3)
        self._check(lib().hf_properties(self.h, nd, _dp(D), _dp(q), _dp(mu)))
        return q, mu

    def core_hamiltonian(self):
        """h = T + V + E . M, the matrix the SCF starts from (HartreeFockAlgorithm::h)."""
        n = self.nbf
        h = np.zeros((n, n))
        self._check(lib().hf_core_hamiltonian(self.h, _dp(h)))
        return h

    def one_electron(self):
        n = self.nbf
        S, T, V = (np.zeros((n, n)) for _ in range(3))
        self._check(lib().hf_one_electron(self.h, _dp(S), _dp(T), _dp(V)))
        return S, T, V

    def scf(self, unrestricted=False, want_density=False, check=True, **kw):
        p = ScfParams()
        lib().hf_scf_defaults(ctypes.byref(p))
        p.unrestricted = 1 if unrestricted else 0
        for k, v in kw.items():
            if not hasattr(p, k):
                raise TypeError("unknown SCF parameter %s" % k)
            setattr(p, k, v)
        r = ScfResult()
        n = self.nbf
        D = np.zeros((2 if unrestricted else 1, n, n)) if want_density else None
        rc = lib().hf_scf(self.h, ctypes.byref(p), ctypes.byref(r), _dp(D) if D is not None else None)
        if check:
            self._check(rc)
        return (r, D) if want_density else r

    def scf_begin(self, unrestricted=False, **kw):
        """Start an SCF driven step by step (HartreeFockAlgorithm::Step)."""
        p = ScfParams()
        lib().hf_scf_defaults(ctypes.byref(p))
        p.unrestricted = 1 if unrestricted else 0
        for k, v in kw.items():
            if not hasattr(p, k):
                raise TypeError("unknown SCF parameter %s" % k)
            setattr(p, k, v)
        self._check(lib().hf_scf_begin(self.h, ctypes.byref(p)))
        self._stepping = 2 if unrestricted else 1
        return self

    def scf_step(self, it):
        """One Step(it): returns (energy, ||D_new - D||_F, DIIS error estimate)."""
        e, r, d = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(lib().hf_scf_step(self.h, it, ctypes.byref(e), ctypes.byref(r), ctypes.byref(d)))
        return e.value, r.value, d.value

    def scf_end(self):
        """End a stepped SCF: (result, densities)."""
        n = self.nbf
        r = ScfResult()
        D = np.zeros((self._stepping, n, n))
        self._check(lib().hf_scf_end(self.h, ctypes.byref(r), _dp(D)))
        return r, D

    def kernel_stats(self):
        arr = (KernelStat * 128)()
        k = lib().hf_kernel_stats(self.h, arr, 128)
        return [arr[i] for i in range(k)]

    def reset_kernel_stats(self):
        self._check(lib().hf_kernel_stats_reset(self.h))


def scf_batch(engines, unrestricted=False, **kw):
    """Lockstep SCF of several molecules (hf_scf_batch): engines is a list of Engine objects, each with its
    molecule set; unrestricted is one flag or one per engine; kw are SCF parameters for every molecule.
    Returns [(status, ScfResult)] in the engines' order (status 0 = converged)."""
    count = len(engines)
    flags = unrestricted if isinstance(unrestricted, (list, tuple)) else [unrestricted] * count
    params = (ScfParams * count)()
    for i in range(count):
        lib().hf_scf_defaults(ctypes.byref(params[i]))
        params[i].unrestricted = 1 if flags[i] else 0
        for k, v in kw.items():
            if not hasattr(params[i], k):
                raise TypeError("unknown SCF parameter %s" % k)
            setattr(params[i], k, v)
    hs = (ctypes.c_void_p * count)(*[e.h for e in engines])
    res = (ScfResult * count)()
    st = (ctypes.c_int * count)()
    lib().hf_scf_batch(hs, count, params, res, st)
    return [(st[i], res[i]) for i in range(count)]


class Server:
    """Continuous-batching SCF server on one GPU (hf_serve_* in include/hfgpu.h).

    Server(slots=32, hbm_budget=0, options={...context options...}, scf={...SCF parameters...}, lockstep=True,
           prep_threads=8): lockstep steps the admitted requests together (one launch per phase for all the
           requests of a batch; True: the library's default of two batch threads, an int: that many);
           lockstep=False runs one worker thread per slot.
    submit(basis_file, atoms, unrestricted=False, tag=0) -> id;  poll(max, wait_ms) -> [ServeResult]
    """

    def __init__(self, slots=None, hbm_budget=0.0, options=None, scf=None, lockstep=True, prep_threads=None):
        L = lib()
        o = ServeOptions()
        L.hf_serve_default_options(ctypes.byref(o))
        if slots is not None:
            o.slots = slots
        o.hbm_budget = hbm_budget
        if lockstep is False or lockstep == 0:
            o.lockstep = 0
        elif lockstep is not True:
            o.lockstep = int(lockstep)
        if prep_threads is not None:
            o.prep_threads = prep_threads
        for k, v in (options or {}).items():
            if not hasattr(o.ctx, k):
                raise TypeError("unknown option %s" % k)
            setattr(o.ctx, k, v)
        for k, v in (scf or {}).items():
            if not hasattr(o.scf, k):
                raise TypeError("unknown SCF parameter %s" % k)
            setattr(o.scf, k, v)
        h = ctypes.c_void_p()
        rc = L.hf_serve_create(ctypes.byref(o), ctypes.byref(h))
        if rc != 0:
            raise HFError(rc, "hf_serve_create failed (is a GPU visible?)")
        self.h = h

    def submit(self, basis_file, atoms, unrestricted=False, tag=0):
        z = np.ascontiguousarray([a[0] for a in atoms], dtype=np.int32)
        xyz = np.ascontiguousarray([c for a in atoms for c in a[1]], dtype=np.float64)
        rid = ctypes.c_uint64()
        rc = lib().hf_serve_submit(self.h, basis_file.encode(), len(atoms), _ip(z), _dp(xyz), 1 if unrestricted else 0,
                                   tag, ctypes.byref(rid))
        if rc != 0:
            raise HFError(rc, "hf_serve_submit failed" + (" (footprint above the HBM budget: rejected)"
                                                          if rc == -2 else ""))
        return rid.value

    def poll(self, max_results=64, wait_ms=100):
        arr = (ServeResult * max_results)()
        k = lib().hf_serve_poll(self.h, arr, max_results, wait_ms)
        return [arr[i] for i in range(k)]

    def pending(self):
        return lib().hf_serve_pending(self.h)

    def kernel_stats(self, reset=False):
        """Per-class kernel statistics of the requests finished so far, summed over the slots."""
        arr = (KernelStat * 128)()
        k = lib().hf_serve_kernel_stats(self.h, arr, 128, 1 if reset else 0)
        return [arr[i] for i in range(k)]

    def close(self):
        if getattr(self, "h", None):
            lib().hf_serve_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serve_footprint(nbf, unrestricted=False):
    """Device bytes the server's admission rule assumes for one SCF on nbf basis functions."""
    return lib().hf_serve_footprint(nbf, 1 if unrestricted else 0)
