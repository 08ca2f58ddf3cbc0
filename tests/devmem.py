"""Device buffers for the GPU tests of the device-pointer entry points, from the HIP runtime libhfgpu.so itself
links (libamdhip64.so.7).  (torch bundles a second HIP runtime; loading both into one process leaves the second
without devices, so the tests do not use torch for device memory.)"""
import ctypes

import numpy as np

_HIP = None
H2D, D2H = 1, 2


def hip():
    global _HIP
    if _HIP is None:
        import hartreefock_amd as hf
        hf.lib()                                  # loads libamdhip64.so.7 as the library's dependency
        L = ctypes.CDLL("libamdhip64.so.7")
        vp = ctypes.c_void_p
        L.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        L.hipFree.argtypes = [vp]
        L.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        L.hipDeviceSynchronize.argtypes = []
        L.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        L.hipStreamDestroy.argtypes = [vp]
        _HIP = L
    return _HIP


def sync():
    assert hip().hipDeviceSynchronize() == 0


class DevArray:
    """A device buffer holding an array of the given shape and dtype (float64 by default)."""

    def __init__(self, shape, dtype=np.float64, fill=None):
        self.shape, self.dtype = tuple(np.atleast_1d(shape)), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 8)) == 0
        self.ptr = p.value
        if fill is not None:
            self.upload(np.full(self.shape, fill, self.dtype))

    @classmethod
    def of(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        sync()
        assert hip().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, H2D) == 0

    def get(self):
        out = np.empty(self.shape, self.dtype)
        sync()
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            sync()
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """A caller-owned HIP stream (default priority)."""

    def __init__(self):
        s = ctypes.c_void_p()
        assert hip().hipStreamCreate(ctypes.byref(s)) == 0
        self.handle = s.value

    def close(self):
        if self.handle:
            sync()
            hip().hipStreamDestroy(self.handle)
            self.handle = None
