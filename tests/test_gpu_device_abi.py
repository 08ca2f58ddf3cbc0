"""The device-pointer entry points of the C ABI (hf_eri_packed_device, hf_eri_pairs_device,
hf_build_jk_incore_device, hf_build_jk_device), called with device buffers the caller owns, on the library's
stream and on a caller's stream: the same results as the host entry points."""
import numpy as np
import pytest

from conftest import basis
from devmem import DevArray, Stream, sync
from hartreefock_amd import geometry as g

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("on_caller_stream", [False, True], ids=["library_stream", "caller_stream"])
def test_device_entry_points_match_host(gpu_engine_factory, on_caller_stream):
    import hartreefock_amd as hf
    atoms = g.alkane_zigzag(3, seed=9)
    eng = gpu_engine_factory(schwarz_threshold=1e-12).set_molecule(basis("6-31g_st_.1.nw"), atoms)
    n = eng.nbf
    s = Stream() if on_caller_stream else None
    sp = s.handle if s is not None else None
    try:
        rng = np.random.default_rng(4)
        X = rng.normal(size=(2, n, n)) * 0.1
        D = X + X.transpose(0, 2, 1)
        dD = DevArray.of(D)
        # packed ERIs
        ln = hf.packed_len(n)
        dP = DevArray(ln, fill=0.0)
        eng.eri_packed_device(dP.ptr, ln, stream=sp)
        assert np.abs(dP.get() - eng.eri_packed()).max() <= 1e-15
        # direct J/K
        dJ = DevArray((n, n), fill=0.0)
        dK = DevArray((2, n, n), fill=0.0)
        eng.build_jk_device(dD.ptr, dJ.ptr, dK.ptr, ndens=2, stream=sp)
        J, K = eng.build_jk(D)
        assert np.abs(dJ.get() - J).max() <= 1e-12 * np.abs(J).max()
        assert np.abs(dK.get() - K).max() <= 1e-12 * np.abs(K).max()
        # in-core: the caller's store
        lg = hf.lib().hf_pair_matrix_len(n)
        dG = DevArray(lg, fill=0.0)
        eng.eri_pairs_device(dG.ptr, lg, stream=sp)
        eng.build_jk_incore_device(dG.ptr, dD.ptr, dJ.ptr, dK.ptr, ndens=2, stream=sp)
        sync()
        Ji, Ki = eng.build_jk_incore(D)
        assert np.abs(dJ.get() - Ji).max() <= 1e-12 * np.abs(Ji).max()
        assert np.abs(dK.get() - Ki).max() <= 1e-12 * np.abs(Ki).max()
        assert np.abs(Ji - J).max() <= 1e-11 * np.abs(J).max()
    finally:
        if s is not None:
            s.close()
