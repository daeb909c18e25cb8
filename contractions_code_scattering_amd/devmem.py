"""Device buffers on the library's own HIP runtime (lc_dev_alloc / lc_memcpy_*).

A DeviceArray owns one allocation and remembers shape and dtype; it is created from a numpy
array (copied to HBM) or empty, and copied back with to_numpy().
"""
import ctypes

import numpy as np

from ._capi import check, lib, register

_VP, _LL = ctypes.c_void_p, ctypes.c_longlong
register("lc_dev_alloc", [_LL, ctypes.POINTER(_VP)])
register("lc_dev_free", [_VP])
register("lc_memcpy_h2d", [_VP, _VP, _LL])
register("lc_memcpy_d2h", [_VP, _VP, _LL])
register("lc_synchronize", [])
register("lc_memcpy_d2d", [_VP, _VP, _LL])


class DeviceArray:
    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = _VP()
        check(lib().lc_dev_alloc(self.nbytes, ctypes.byref(p)))
        self.ptr = p.value
        self._lib = lib()

    @classmethod
    def from_numpy(cls, a, dtype=None):
        a = np.ascontiguousarray(a if dtype is None else a.astype(dtype))
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            check(lib().lc_memcpy_h2d(d.ptr, a.ctypes.data, d.nbytes))
        return d

    def to_numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib().lc_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr:
            self._lib.lc_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def data_ptr(self):
        return self.ptr


def to_device(a, dtype=None):
    return DeviceArray.from_numpy(a, dtype)


def synchronize():
    check(lib().lc_synchronize())


def copy(dst, src):
    """Device-to-device copy of src into dst (same byte size)."""
    if dst.nbytes != src.nbytes:
        raise ValueError("size mismatch")
    check(lib().lc_memcpy_d2d(dst.ptr, src.ptr, src.nbytes))


register("lc_read_prop_device", [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP])


def read_perambulator(path, nt, nv, dtype=np.complex128):
    """Perambulator file (the reference's readPropH format) -> DeviceArray [nt][nt][4][4][nv][nv],
    permuted on the GPU (lc_read_prop_device); dtype complex128 or complex64."""
    dtype = np.dtype(dtype)
    if dtype not in (np.dtype(np.complex128), np.dtype(np.complex64)):
        raise TypeError("perambulator dtype must be complex128 or complex64")
    out = DeviceArray((nt, nt, 4, 4, nv, nv), dtype)
    prec = 1 if dtype == np.complex128 else 0
    check(lib().lc_read_prop_device(str(path).encode(), nt, nv, prec, out.ptr))
    return out
