"""Device buffers for tests that hand device addresses to the C ABI.

Uses the HIP runtime the library itself loaded (libamdhip64.so.7 under
/opt/rocm), through ctypes; no second runtime enters the process.
"""
import ctypes as C

import numpy as np

_hip = None
H2D, D2H = 1, 2


def hip():
    global _hip
    if _hip is None:
        import eph_amd
        eph_amd.lib()  # the library's own runtime is loaded first
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DeviceArray:
    """A float64 array copied into device memory (freed on close)."""

    def __init__(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.n = a.size
        self.ptr = C.c_void_p()
        nbytes = max(a.nbytes, 8)
        rc = hip().hipMalloc(C.byref(self.ptr), nbytes)
        assert rc == 0, f"hipMalloc failed ({rc})"
        if a.nbytes:
            rc = hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D)
            assert rc == 0, f"hipMemcpy failed ({rc})"

    @property
    def addr(self):
        return self.ptr.value

    def close(self):
        if self.ptr and self.ptr.value:
            hip().hipFree(self.ptr)
        self.ptr = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
