"""Device buffers through the HIP runtime (ctypes): allocation and copies for
the `_dev` entry points and the legacy ABI's device-pointer convention.
Plumbing only; no compute happens here."""
from __future__ import annotations

import ctypes as C

import numpy as np

_H2D, _D2H = 1, 2


class Hip:
    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipDeviceSynchronize.argtypes = []

    def malloc(self, n: int) -> int:
        p = C.c_void_p()
        if self.L.hipMalloc(C.byref(p), C.c_size_t(max(int(n), 1))) != 0:
            raise MemoryError(f"hipMalloc of {n} bytes failed")
        return p.value

    def h2d(self, d: int, h: np.ndarray):
        h = np.ascontiguousarray(h)
        if self.L.hipMemcpy(C.c_void_p(d), h.ctypes.data, h.nbytes, _H2D) != 0:
            raise RuntimeError("hipMemcpy host->device failed")

    def d2h(self, h: np.ndarray, d: int):
        assert h.flags.c_contiguous
        if self.L.hipMemcpy(h.ctypes.data, C.c_void_p(d), h.nbytes, _D2H) != 0:
            raise RuntimeError("hipMemcpy device->host failed")

    def to_device(self, h: np.ndarray) -> int:
        d = self.malloc(h.nbytes)
        self.h2d(d, h)
        return d

    def memset(self, d: int, v: int, n: int):
        if self.L.hipMemset(C.c_void_p(d), int(v), C.c_size_t(n)) != 0:
            raise RuntimeError("hipMemset failed")

    def free(self, d: int):
        self.L.hipFree(C.c_void_p(d))

    def sync(self):
        self.L.hipDeviceSynchronize()
