"""Device buffers for tests through the HIP runtime (ctypes on libamdhip64.so)."""

import ctypes as C
import ctypes.util
import os

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        path = "/opt/rocm/lib/libamdhip64.so"
        if not os.path.exists(path):
            path = ctypes.util.find_library("amdhip64")
        L = C.CDLL(path)
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipDeviceSynchronize.argtypes = []
        L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        _hip = L
    return _hip


H2D, D2H = 1, 2


class DeviceArray:
    """A device copy of a numpy array (or zeros of a shape / dtype)."""

    def __init__(self, host=None, shape=None, dtype=None):
        if host is not None:
            host = np.ascontiguousarray(host)
            shape, dtype = host.shape, host.dtype
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 8)) == 0
        if host is not None:
            assert hip().hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, H2D) == 0
        else:
            assert hip().hipMemset(self.ptr, 0, max(self.nbytes, 8)) == 0

    def get(self):
        out = np.empty(self.shape, dtype=self.dtype)
        assert hip().hipDeviceSynchronize() == 0
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) == 0
        return out

    def __del__(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()
