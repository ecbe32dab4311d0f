"""Page-locked host arrays (hipHostMalloc), for host-to-device uploads that
run at DMA speed.  Falls back to an ordinary array when the runtime refuses."""
from __future__ import annotations

import ctypes

import numpy as np

_HIP = None


def _hip():
    global _HIP
    if _HIP is None:
        lib = ctypes.CDLL("libamdhip64.so")     # already loaded by libsinfvm_hip.so
        lib.hipHostMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_uint]
        lib.hipHostMalloc.restype = ctypes.c_int
        lib.hipHostFree.argtypes = [ctypes.c_void_p]
        lib.hipHostFree.restype = ctypes.c_int
        _HIP = lib
    return _HIP


class PinnedArray:
    """A Fortran-ordered float64 array in page-locked host memory; free() (or
    garbage collection) releases it."""

    def __init__(self, shape):
        self.shape = tuple(int(s) for s in shape)
        nbytes = 8 * int(np.prod(self.shape))
        self._ptr = ctypes.c_void_p()
        try:
            rc = _hip().hipHostMalloc(ctypes.byref(self._ptr), nbytes, 0)
        except OSError:
            rc = -1
        if rc != 0 or not self._ptr.value:
            self._ptr = ctypes.c_void_p()
            self.array = np.empty(self.shape, order="F")
            self.pinned = False
            return
        buf = (ctypes.c_char * nbytes).from_address(self._ptr.value)
        self.array = np.frombuffer(buf, dtype=np.float64).reshape(self.shape, order="F")
        self.pinned = True

    def free(self) -> None:
        if self._ptr.value:
            self.array = None
            _hip().hipHostFree(self._ptr)
            self._ptr = ctypes.c_void_p()

    def __del__(self):
        self.free()
