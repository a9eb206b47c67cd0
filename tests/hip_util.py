"""Minimal device-memory helper for tests: calls the HIP runtime the backend
library itself loaded (libamdhip64.so.7 from /opt/rocm), so no second HIP
runtime is brought into the process."""

import ctypes as C

import numpy as np

from tessera_amd.mpm import load_library

_HIP = None


def hip():
    global _HIP
    if _HIP is None:
        load_library()  # the backend pulls in its HIP runtime first
        h = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        h.hipDeviceSynchronize.argtypes = []
        _HIP = h
    return _HIP


class DeviceBuffer:
    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        assert hip().hipMalloc(C.byref(p), self.nbytes) == 0
        assert hip().hipMemset(p, 0, self.nbytes) == 0
        self.ptr = p.value

    def read(self, dtype) -> np.ndarray:
        out = np.zeros(self.nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        assert hip().hipDeviceSynchronize() == 0
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) == 0  # device to host
        return out

    def __del__(self):
        try:
            hip().hipFree(self.ptr)
        except Exception:
            pass
