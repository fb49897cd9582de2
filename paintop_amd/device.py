"""Device memory for callers of the device-resident entry points.

Bound with ctypes to the same HIP runtime (libamdhip64.so.7 of /opt/rocm) that
libpaintop_hip.so links, so a process holds exactly one HIP runtime. (PyTorch
wheels bundle their own HIP/HSA runtime; initialising both in one process
does not work, so nothing on the GPU path here goes through torch.)
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import hip as _hip

_rt = None
H2D, D2H, D2D = 1, 2, 3


def runtime():
    """The HIP runtime already loaded by libpaintop_hip.so."""
    global _rt
    if _rt is None:
        _hip.lib()  # loads libpaintop_hip.so and, through it, its HIP runtime
        rt = ctypes.CDLL("libamdhip64.so.7")
        rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        rt.hipFree.argtypes = [ctypes.c_void_p]
        rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        rt.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        rt.hipDeviceSynchronize.argtypes = []
        rt.hipGetErrorString.argtypes = [ctypes.c_int]
        rt.hipGetErrorString.restype = ctypes.c_char_p
        _rt = rt
    return _rt


def _ok(code: int, what: str):
    if code != 0:
        raise RuntimeError(f"{what} failed: {runtime().hipGetErrorString(code).decode()}")


def synchronize():
    _ok(runtime().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A contiguous device buffer with a numpy-like shape and dtype."""

    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _ok(runtime().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = int(p.value)

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            _ok(runtime().hipMemcpy(d.ptr, a.ctypes.data, d.nbytes, H2D), "hipMemcpy H2D")
        return d

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            _ok(runtime().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def zero(self) -> "DeviceArray":
        _ok(runtime().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")
        return self

    def copy_from(self, other: "DeviceArray") -> "DeviceArray":
        assert other.nbytes == self.nbytes
        _ok(runtime().hipMemcpy(self.ptr, other.ptr, self.nbytes, D2D), "hipMemcpy D2D")
        return self

    def free(self):
        if getattr(self, "ptr", 0):
            runtime().hipFree(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
