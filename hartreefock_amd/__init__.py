"""hartreefock_amd — MI355X (gfx950) ERI / J-K Fock-build engine with a device-resident SCF.

Python binding of the C ABI in include/hfgpu.h (libhfgpu.so, built in-tree by
`python -m hartreefock_amd.build`).  There is no CPU fallback: every call runs the
HIP library, and importing the binding raises if libhfgpu.so is missing.
"""
from ._lib import (Engine, HFError, KernelStat, Options, ScfParams, ScfResult, Server, ServeResult, lib,
                   library_path, packed_index, packed_len, scf_batch, serve_footprint, EXPORTED_SYMBOLS)

__all__ = ["Engine", "HFError", "KernelStat", "Options", "ScfParams", "ScfResult", "Server", "ServeResult", "lib",
           "library_path", "packed_index", "packed_len", "scf_batch", "serve_footprint", "EXPORTED_SYMBOLS",
           "basis_path"]


def basis_path(name):
    """Path of a basis-set file shipped in hartreefock_amd/basis (e.g. '6-31g_st_.1.nw')."""
    import os
    p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis", name)
    if not os.path.exists(p):
        raise FileNotFoundError(p)
    return p
