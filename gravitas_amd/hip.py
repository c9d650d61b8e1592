"""Device memory through the HIP runtime that libgravitas_hip.so links.

torch wheels bundle their own HIP/HSA runtime; two runtimes in one process
contend for the GPU, so the backend's buffers are plain hipMalloc'd arrays
managed here (ctypes over libamdhip64, which the backend library has already
loaded, so both share one runtime instance).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

_rt = None
H2D, D2H, D2D = 1, 2, 3


def rt():
    global _rt
    if _rt is None:
        _lib.load()  # the backend pulls in /opt/rocm's libamdhip64 first
        R = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
        R.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        R.hipFree.argtypes = [ctypes.c_void_p]
        R.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        R.hipDeviceSynchronize.argtypes = []
        R.hipGetErrorString.argtypes = [ctypes.c_int]
        R.hipGetErrorString.restype = ctypes.c_char_p
        _rt = R
    return _rt


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {rt().hipGetErrorString(rc).decode()}")


def synchronize():
    _check(rt().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A 1-D or 2-D array in HBM with numpy-style dtype and shape."""

    def __init__(self, shape, dtype):
        self.shape = (shape,) if np.isscalar(shape) else tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _check(rt().hipMalloc(ctypes.byref(p), max(self.nbytes, 256)), "hipMalloc")
        self._p = p

    @property
    def ptr(self):
        return self._p

    def view(self, start, count):
        """Non-owning view of elements [start, start+count) of a 1-D array."""
        v = DeviceArray.__new__(DeviceArray)
        v.shape = (count,)
        v.dtype = self.dtype
        v.nbytes = count * self.dtype.itemsize
        v._p = ctypes.c_void_p((self._p.value or 0) + start * self.dtype.itemsize)
        v._owner = self
        v._view = True
        return v

    def __len__(self):
        return self.shape[0]

    def __del__(self):
        try:
            if self._p and not getattr(self, "_view", False):
                rt().hipFree(self._p)
            self._p = None
        except Exception:
            pass

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        if self.nbytes:
            _check(rt().hipMemcpy(self._p, a.ctypes.data_as(ctypes.c_void_p), self.nbytes, H2D), "hipMemcpy H2D")

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            _check(rt().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self._p, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def fill(self, value):
        self.upload(np.full(self.shape, value, self.dtype))
