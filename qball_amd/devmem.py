"""Device buffers through the HIP runtime that libqbh.so is linked against, without PyTorch.

A C++ host (qball) links libqbh.so with the ROCm libraries of the system.  A Python process
that imports torch first binds libqbh.so to the HIP runtime and BLAS libraries bundled with
torch instead (they carry the same sonames), so a process that wants the library exactly
as a C++ host runs it allocates its buffers here and never imports torch.
"""
import ctypes

import numpy as np

_HIP = None
H2D, D2H, D2D = 1, 2, 3


def hip():
    """The HIP runtime already loaded with libqbh.so (looked up by soname)."""
    global _HIP
    if _HIP is None:
        from . import capi
        capi.lib()
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        h.hipSetDevice.argtypes = [ctypes.c_int]
        h.hipGetErrorString.restype = ctypes.c_char_p
        h.hipGetErrorString.argtypes = [ctypes.c_int]
        _HIP = h
    return _HIP


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


def synchronize():
    _check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """nbytes of device memory on the current device; freed by close() or on collection."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self._p = p.value

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        d.upload(a)
        return d

    @property
    def ptr(self):
        return self._p

    def data_ptr(self):
        return self._p

    def upload(self, a, offset=0):
        """Copies host array a to byte offset `offset` (waits for the device first)."""
        a = np.ascontiguousarray(a)
        if offset + a.nbytes > self.nbytes:
            raise ValueError("upload past the end of the buffer")
        synchronize()
        _check(hip().hipMemcpy(self._p + offset, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")

    def to_numpy(self, dtype=np.float64, count=None, offset=0):
        """count elements of dtype from byte offset `offset` (after the device drains)."""
        dt = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - offset) // dt.itemsize
        if offset + count * dt.itemsize > self.nbytes:
            raise ValueError("read past the end of the buffer")
        out = np.empty(count, dtype=dt)
        synchronize()
        _check(hip().hipMemcpy(out.ctypes.data, self._p + offset, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def zero(self):
        synchronize()
        _check(hip().hipMemset(self._p, 0, self.nbytes), "hipMemset")
        synchronize()

    def close(self):
        if self._p:
            synchronize()
            hip().hipFree(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
