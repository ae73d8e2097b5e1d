"""Device buffers for GPU tests through the HIP runtime libdiablo_hip.so
itself links (no second runtime in the process)."""
import ctypes

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        from diablo_amd import channel
        channel.load_library()  # loads /opt/rocm's libamdhip64.so.7 first
        lib = ctypes.CDLL("libamdhip64.so.7")
        lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        lib.hipFree.argtypes = [ctypes.c_void_p]
        lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        lib.hipDeviceSynchronize.argtypes = []
        _hip = lib
    return _hip


class DeviceArray:
    def __init__(self, nbytes):
        self.nbytes = nbytes
        self.ptr = ctypes.c_void_p()
        rc = hip().hipMalloc(ctypes.byref(self.ptr), nbytes)
        assert rc == 0, f"hipMalloc {rc}"

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1) == 0

    def get(self, dtype, shape):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        assert hip().hipDeviceSynchronize() == 0
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None
