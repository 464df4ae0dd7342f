"""Device buffers for tests of entry points that take device pointers: hipMalloc / hipMemcpy /
hipFree of the same HIP runtime the library links (libamdhip64), through ctypes."""
import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


H2D, D2H = 1, 2


class DeviceArray:
    """A device copy of a numpy array (shape and dtype kept for the way back)."""

    def __init__(self, a=None, shape=None, dtype=None):
        if a is not None:
            a = np.ascontiguousarray(a)
            shape, dtype = a.shape, a.dtype
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        rc = hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 16))
        assert rc == 0, f"hipMalloc failed: {rc}"
        if a is not None and self.nbytes:
            assert hip().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, H2D) == 0

    def get(self):
        out = np.empty(self.shape, self.dtype)
        hip().hipDeviceSynchronize()
        if self.nbytes:
            assert hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) == 0
        return out

    def data_ptr(self):
        return self.ptr.value

    def __del__(self):
        try:
            if self.ptr:
                hip().hipFree(self.ptr)
                self.ptr = C.c_void_p()
        except Exception:
            pass
