"""The device-pointer entry points of the C ABI (hf_eri_packed_device, hf_eri_pairs_device,
hf_build_jk_incore_device, hf_build_jk_device), called with device buffers the caller owns, on the library's
stream and on a caller's stream: the same results as the host entry points."""
import numpy as np
import pytest

from conftest import basis
from hartreefock_amd import geometry as g

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available()
    return torch


@pytest.mark.parametrize("on_caller_stream", [False, True], ids=["library_stream", "caller_stream"])
def test_device_entry_points_match_host(gpu_engine_factory, on_caller_stream):
    torch = _torch()
    import hartreefock_amd as hf
    atoms = g.alkane_zigzag(3, seed=9)
    eng = gpu_engine_factory(schwarz_threshold=1e-12).set_molecule(basis("6-31g_st_.1.nw"), atoms)
    n = eng.nbf
    s = torch.cuda.Stream() if on_caller_stream else None
    sp = s.cuda_stream if s is not None else None
    rng = np.random.default_rng(4)
    X = rng.normal(size=(2, n, n)) * 0.1
    D = X + X.transpose(0, 2, 1)
    dD = torch.tensor(D, device="cuda")
    # packed ERIs
    ln = hf.packed_len(n)
    dP = torch.zeros(ln, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    eng.eri_packed_device(dP.data_ptr(), ln, stream=sp)
    torch.cuda.synchronize()
    assert np.abs(dP.cpu().numpy() - eng.eri_packed()).max() <= 1e-15
    # direct J/K
    dJ = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    dK = torch.zeros((2, n, n), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    eng.build_jk_device(dD.data_ptr(), dJ.data_ptr(), dK.data_ptr(), ndens=2, stream=sp)
    torch.cuda.synchronize()
    J, K = eng.build_jk(D)
    assert np.abs(dJ.cpu().numpy() - J).max() <= 1e-12 * np.abs(J).max()
    assert np.abs(dK.cpu().numpy() - K).max() <= 1e-12 * np.abs(K).max()
    # in-core: the caller's store
    lg = hf.lib().hf_pair_matrix_len(n)
    dG = torch.zeros(lg, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    eng.eri_pairs_device(dG.data_ptr(), lg, stream=sp)
    eng.build_jk_incore_device(dG.data_ptr(), dD.data_ptr(), dJ.data_ptr(), dK.data_ptr(), ndens=2, stream=sp)
    torch.cuda.synchronize()
    Ji, Ki = eng.build_jk_incore(D)
    assert np.abs(dJ.cpu().numpy() - Ji).max() <= 1e-12 * np.abs(Ji).max()
    assert np.abs(dK.cpu().numpy() - Ki).max() <= 1e-12 * np.abs(Ki).max()
    assert np.abs(Ji - J).max() <= 1e-11 * np.abs(J).max()
