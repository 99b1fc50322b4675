"""Device buffers and streams for tests that call device-pointer entry points,
through the HIP runtime the library itself links (libamdhip64.so.7 from
/opt/rocm), so one runtime instance serves both.  Test plumbing only."""
from __future__ import annotations

import ctypes

import numpy as np

_H2D, _D2H = 1, 2


def _rt():
    from tremor_amd import capi
    capi.lib()  # the library's own dependency is loaded first; the same soname then resolves to it
    r = ctypes.CDLL("libamdhip64.so.7")
    r.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    r.hipFree.argtypes = [ctypes.c_void_p]
    r.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    r.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    r.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    r.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    r.hipDeviceSynchronize.argtypes = []
    return r


class DeviceArray:
    """A device copy of a numpy array (or an uninitialised one of a shape)."""

    def __init__(self, host: np.ndarray):
        self._rt = _rt()
        self.dtype = host.dtype
        self.size = host.size
        self.nbytes = max(host.nbytes, 1)
        p = ctypes.c_void_p()
        if self._rt.hipMalloc(ctypes.byref(p), self.nbytes) != 0:
            raise MemoryError("hipMalloc failed")
        self.ptr = p.value
        h = np.ascontiguousarray(host)
        if self._rt.hipMemcpy(self.ptr, h.ctypes.data, h.nbytes, _H2D) != 0:
            raise RuntimeError("hipMemcpy to device failed")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.size, dtype=self.dtype)
        if self._rt.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, _D2H) != 0:
            raise RuntimeError("hipMemcpy to host failed")
        return out

    def free(self):
        if self.ptr:
            self._rt.hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()


class Stream:
    def __init__(self):
        self._rt = _rt()
        s = ctypes.c_void_p()
        if self._rt.hipStreamCreate(ctypes.byref(s)) != 0:
            raise RuntimeError("hipStreamCreate failed")
        self.handle = s.value

    def synchronize(self):
        if self._rt.hipStreamSynchronize(self.handle) != 0:
            raise RuntimeError("hipStreamSynchronize failed")

    def destroy(self):
        if self.handle:
            self._rt.hipStreamDestroy(self.handle)
            self.handle = None


def device_synchronize():
    if _rt().hipDeviceSynchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")
