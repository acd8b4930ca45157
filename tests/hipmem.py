"""Device buffers for tests that call the C ABI with device pointers:
hipMalloc / hipMemcpy / hipFree from the HIP runtime through ctypes."""
import ctypes

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        lib = ctypes.CDLL("libamdhip64.so")
        lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        lib.hipFree.argtypes = [ctypes.c_void_p]
        lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        lib.hipDeviceSynchronize.argtypes = []
        _hip = lib
    return _hip


class DeviceArray:
    """A float64 device buffer with host copies in and out."""

    def __init__(self, host):
        self.host = np.ascontiguousarray(host, dtype=np.float64)
        self.ptr = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.ptr), max(self.host.nbytes, 8)) == 0
        assert hip().hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, 1) == 0   # host -> device

    @property
    def addr(self):
        return self.ptr.value

    def get(self):
        out = np.empty_like(self.host)
        assert hip().hipDeviceSynchronize() == 0
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0   # device -> host
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()
