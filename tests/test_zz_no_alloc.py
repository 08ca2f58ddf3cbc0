"""The allocation-free in-core J/K entry point (hf_build_jk_incore_device_noalloc): its scratch is a static device
array shared by every call on the GPU.  Two calls in a row (the second on a smaller molecule, so stale contents
of the shared array would show) against hf_build_jk_incore_device on the same store and densities; a basis above
the array's limit (n > 512) is refused with HF_E_ARG."""
import numpy as np
import pytest

from conftest import basis
from devmem import DevArray, sync
from hartreefock_amd import geometry as g

pytestmark = pytest.mark.gpu


def _case(eng, ndens, seed):
    import hartreefock_amd as hf
    n = eng.nbf
    lg = hf.lib().hf_pair_matrix_len(n)
    dG = DevArray(lg, fill=0.0)
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(ndens, n, n)) * 0.1
    dD = DevArray.of(X + X.transpose(0, 2, 1))
    J = [DevArray((n, n), fill=7.0) for _ in range(2)]
    K = [DevArray((ndens, n, n), fill=7.0) for _ in range(2)]
    eng.eri_pairs_device(dG.ptr, lg)
    eng.build_jk_incore_device(dG.ptr, dD.ptr, J[0].ptr, K[0].ptr, ndens=ndens)
    eng.build_jk_incore_device(dG.ptr, dD.ptr, J[1].ptr, K[1].ptr, ndens=ndens, no_alloc=True)
    sync()
    J0, J1, K0, K1 = J[0].get(), J[1].get(), K[0].get(), K[1].get()
    print("n %d ndens %d: max|dJ| %.2e max|dK| %.2e" % (n, ndens, np.abs(J1 - J0).max(), np.abs(K1 - K0).max()),
          flush=True)
    assert np.abs(J1 - J0).max() <= 1e-13 * np.abs(J0).max()
    assert np.abs(K1 - K0).max() <= 1e-13 * np.abs(K0).max()


def test_noalloc_matches_incore_device(gpu_engine_factory):
    eng = gpu_engine_factory(schwarz_threshold=1e-12)
    eng.set_molecule(basis("6-31g_st_.1.nw"), g.alkane_zigzag(5, seed=31))      # 99 functions, UHF densities
    _case(eng, 2, 1)
    eng.set_molecule(basis("6-31g_st_.1.nw"), g.alkane_zigzag(2, seed=5))       # 42 functions, then smaller
    _case(eng, 1, 2)


def test_noalloc_rejects_basis_above_limit(gpu_engine_factory):
    import hartreefock_amd as hf
    eng = gpu_engine_factory().set_molecule(basis("6-31g_st__st_.nw"), g.water32_cfg3())     # 800 functions
    x = DevArray(16, fill=0.0)
    with pytest.raises(hf.HFError) as ei:
        eng.build_jk_incore_device(x.ptr, x.ptr, x.ptr, x.ptr, no_alloc=True)
    assert ei.value.code == -1 and "512" in str(ei.value)
