"""Device buffers for Python callers of the C ABI (ctypes over the HIP runtime).

libmpflip.so links the system HIP runtime (libamdhip64.so.7 of /opt/rocm);
a Python wheel of PyTorch may carry a second copy of the runtime, and two
runtimes in one process do not share devices.  Callers that pass device
pointers to the ABI (MPF_DEVICE, mpf_hash_sort_noalloc) allocate them here,
from the same runtime the library uses.
"""
import ctypes

import numpy as np

from . import lib

_H = None
H2D, D2H, D2D = 1, 2, 3


def hip():
    global _H
    if _H is None:
        lib()  # loads libamdhip64.so.7 as a dependency of libmpflip.so
        _H = ctypes.CDLL("libamdhip64.so.7")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        _H.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        _H.hipFree.argtypes = [vp]
        _H.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        _H.hipMemset.argtypes = [vp, ctypes.c_int, sz]
        _H.hipDeviceSynchronize.argtypes = []
        _H.hipGetErrorString.argtypes = [ctypes.c_int]
        _H.hipGetErrorString.restype = ctypes.c_char_p
        _H.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
    return _H


def _check(e, what):
    if e != 0:
        raise RuntimeError("%s: %s" % (what, hip().hipGetErrorString(e).decode()))


def device_count():
    n = ctypes.c_int(0)
    e = hip().hipGetDeviceCount(ctypes.byref(n))
    return n.value if e == 0 else 0


def synchronize():
    _check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A device allocation of `count` elements of `dtype` (freed with the object)."""

    def __init__(self, count, dtype):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = max(1, self.count * self.dtype.itemsize)
        p = ctypes.c_void_p()
        _check(hip().hipMalloc(ctypes.byref(p), self.nbytes), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size <= self.count
        if a.size:
            _check(hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")

    def fill_bytes(self, value):
        _check(hip().hipMemset(self.ptr, int(value), self.nbytes), "hipMemset")

    def download(self, count=None):
        n = self.count if count is None else int(count)
        out = np.empty(n, dtype=self.dtype)
        if n:
            _check(hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def copy_from(self, other, nbytes):
        _check(hip().hipMemcpy(self.ptr, other.ptr, int(nbytes), D2D), "hipMemcpy D2D")

    def data_ptr(self):
        return self.ptr

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
