"""The C-ABI library loads and exports every function include/hfgpu.h declares (CPU only:
no compute call is made here)."""
import ctypes
import os
import re

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
