"""The C-ABI library loads and exports every function include/hfgpu.h declares (CPU only:
no compute call is made here)."""
import ctypes
import os
import re

import pytest

import hartreefock_amd as hf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "hfgpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hf_[a-z_0-9]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(hf.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(hf.library_path())
    for name in declared_functions():
        assert hasattr(L, name), name


def test_pure_helpers_without_gpu():
    assert hf.packed_index(0, 0, 0, 0) == 0
    assert hf.packed_index(1, 0, 0, 0) == 1
    assert hf.packed_index(3, 2, 1, 0) == hf.packed_index(0, 1, 2, 3)
    n = 7
    assert hf.packed_len(n) == hf.packed_index(n, n, n, n) + 1
    # serving admission estimate: 56 (RHF) / 64 (UHF) n x n FP64 matrices + 32 MiB of pair and task data
    assert hf.serve_footprint(200) == 8.0 * 200 * 200 * 56 + 32 * 2 ** 20
    assert hf.serve_footprint(200, unrestricted=True) == 8.0 * 200 * 200 * 64 + 32 * 2 ** 20
    assert hf.serve_footprint(800) > hf.serve_footprint(200)


def test_options_defaults():
    o = hf.Options()
    hf.lib().hf_default_options(ctypes.byref(o))
    assert o.schwarz_threshold == 1e-12 and o.exact_transfer == 0 and o.precision == 0
    p = hf.ScfParams()
    hf.lib().hf_scf_defaults(ctypes.byref(p))
    assert (p.alpha, p.init_guess, p.use_diis, p.normal_iter_after_diis) == (0.75, 0.75, 1, 500)


def test_cpp_dropin_header_compiles(tmp_path):
    """The C++ drop-in header (reference class names over the C ABI) builds and links."""
    import subprocess
    exe = tmp_path / "dropin_example"
    cmd = ["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "dropin_example.cpp"),
           "-L" + os.path.join(ROOT, "hartreefock_amd"), "-lhfgpu", "-o", str(exe)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr


@pytest.mark.parametrize("bname", ["sto3g.txt", "6-31g_st_.1.nw", "6-31+g_st_.1.nw"])
def test_cpp_basis_model_matches_oracle(tmp_path, bname):
    """The drop-in header's own NWChem reader and data model (Chemistry::Basis::Load into basis.atoms, shells of
    Cartesian components, per-component normalisation, Molecule::Init IDs) against the oracle's molecule: the same
    basis-function order, (l, m, n), atoms, exponents and coefficients, and normalisation factors
    (HartreeFock/GaussianOrbital.cpp:25-30)."""
    import subprocess
    import numpy as np
    from oracle import refcpu
    from hartreefock_amd import geometry as g
    exe = tmp_path / "basis_model"
    cmd = ["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "basis_model.cpp"),
           "-L" + os.path.join(ROOT, "hartreefock_amd"), "-lhfgpu", "-o", str(exe)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    atoms = g.alkane_zigzag(2, seed=5)
    args = [str(exe), os.path.join(ROOT, "hartreefock_amd", "basis", bname)]
    for z, xyz in atoms:
        args += [str(z)] + ["%.17g" % c for c in xyz]
    env = dict(os.environ, LD_LIBRARY_PATH=os.path.join(ROOT, "hartreefock_amd") + ":" +
               os.environ.get("LD_LIBRARY_PATH", ""))
    out = subprocess.run(args, capture_output=True, text=True, env=env, timeout=60)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    head = lines[0].split()
    mol = refcpu.Molecule(os.path.join(ROOT, "hartreefock_amd", "basis", bname), atoms)
    L, lmn, at, _ = mol.bf_info()
    assert int(head[1]) == mol.nbf and (int(head[3]), int(head[4])) == mol.electrons()
    bfs = [l.split() for l in lines[1:]]
    assert len(bfs) == mol.nbf
    for k, b in enumerate(bfs):
        assert int(b[1]) == k and int(b[2]) == at[k] and int(b[3]) == L[k]
        assert (int(b[4]), int(b[5]), int(b[6])) == tuple(lmn[k])
        a, c, nrm = float(b[7]), float(b[8]), float(b[9])
        l, m, n = lmn[k]
        dfact = lambda x: float(np.prod(np.arange(x, 0, -2))) if x > 0 else 1.0
        ref = (2 * a / np.pi) ** 0.75 * (4 * a) ** ((l + m + n) / 2.0) / np.sqrt(
            dfact(2 * l - 1) * dfact(2 * m - 1) * dfact(2 * n - 1))
        assert abs(nrm - ref) <= 1e-13 * ref and a > 0 and c != 0
        assert abs(float(b[10]) - atoms[at[k]][1][2]) <= 1e-15     # shell centres set by Init()
