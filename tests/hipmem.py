"""Device buffers for tests that pass device pointers to the library,
allocated through the HIP runtime the library itself is bound to.  torch
bundles its own libamdhip64 with the same soname: imported first, it is the
runtime the library binds to; imported after the library, it loads a second
runtime, and once one runtime has opened the GPU the other sees none.  So
the buffers come from the object the dynamic linker bound for the library's
libamdhip64.so.7, which dlopen of that soname returns."""
import ctypes

import numpy as np

_hip = None


def runtime():
    global _hip
    if _hip is None:
        from sqaod_amd.rocm import _cext  # noqa: F401  (loads libsqaod_rocm.so and binds its runtime)
        _hip = ctypes.CDLL("libamdhip64.so.7")
    return _hip


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class DeviceArray:
    """n elements of dtype in device memory."""

    def __init__(self, n, dtype):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.ptr = ctypes.c_void_p()
        _check(runtime().hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(1, self.n * self.dtype.itemsize))),
               "hipMalloc")

    @classmethod
    def of(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        _check(runtime().hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1),
               "hipMemcpy H2D")

    def fill_bytes(self, value):
        _check(runtime().hipMemset(self.ptr, ctypes.c_int(value), ctypes.c_size_t(self.n * self.dtype.itemsize)),
               "hipMemset")

    def download(self):
        out = np.empty(self.n, self.dtype)
        _check(runtime().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(out.nbytes), 2),
               "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            runtime().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    _check(runtime().hipDeviceSynchronize(), "hipDeviceSynchronize")
