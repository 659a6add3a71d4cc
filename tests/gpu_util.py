"""Helpers shared by the GPU tests (not collected: no test_ prefix)."""
import sys

import numpy as np


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def random_density(n, seed, scale=0.3):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) * scale
    P = A @ A.T / n
    return P


def packed(F):
    return F[np.tril_indices(F.shape[0])].copy()


class DeviceArray:
    """A float64 device buffer from the HIP runtime the library links
    (libamdhip64.so.7, through ctypes): tests that call the device entry
    points get device memory without loading a second runtime."""

    _hip = None

    @classmethod
    def hip(cls):
        import ctypes as C

        if cls._hip is None:
            h = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, host):
        import ctypes as C

        a = np.ascontiguousarray(host, dtype=np.float64)
        self.n = a.size
        self.ptr = C.c_void_p()
        assert self.hip().hipMalloc(C.byref(self.ptr), max(8, 8 * self.n)) == 0, "hipMalloc failed"
        assert self.hip().hipMemcpy(self.ptr, a.ctypes.data, 8 * self.n, 1) == 0  # hipMemcpyHostToDevice

    @property
    def addr(self) -> int:
        return self.ptr.value

    def get(self) -> np.ndarray:
        out = np.empty(self.n)
        assert self.hip().hipDeviceSynchronize() == 0
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, 8 * self.n, 2) == 0  # hipMemcpyDeviceToHost
        return out

    def free(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = None
