"""Torch-facing wrappers of the sink-row launchers of the C-ABI.

Operands are DeviceArrays (complex128 -> LC_DOUBLE, complex64 -> LC_SINGLE) in the reference's
device layouts; results come back as numpy arrays in the layout the reference launchers write.
Shapes are checked on the host before any launch.
"""
import ctypes

import numpy as np

from ._capi import DOUBLE, SINGLE, check, lib, register

register("lc_loop_unity_noalloc", [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_void_p])


def init(device=0):
    if not lib().lc_is_init():
        check(lib().lc_init(device))


def finalize():
    if lib().lc_is_init():
        check(lib().lc_finalize())


def _prec(t):
    if t.dtype == np.complex128:
        return DOUBLE, np.complex128
    if t.dtype == np.complex64:
        return SINGLE, np.complex64
    raise TypeError("operands must be complex64 or complex128, got %s" % t.dtype)


def _p(t, shape, dtype):
    from .devmem import DeviceArray
    if not isinstance(t, DeviceArray):
        raise TypeError("operand must be a DeviceArray")
    if tuple(t.shape) != tuple(shape):
        raise ValueError("operand shape %s, expected %s" % (tuple(t.shape), tuple(shape)))
    if t.dtype != dtype:
        raise TypeError("mixed precisions")
    return ctypes.c_void_p(t.ptr)


def _out(shape, npdt):
    o = np.zeros(shape, npdt)
    return o, ctypes.c_void_p(o.ctypes.data)


def _dims(row):
    if len(row.shape) != 5 or row.shape[1:3] != (4, 4) or row.shape[3] != row.shape[4]:
        raise ValueError("row must be [Nt][4][4][Nv][Nv]")
    return row.shape[0], row.shape[-1]


def rho_row(row, mom5, tf):
    nt, nv = _dims(row)
    prec, npdt = _prec(row)
    out, po = _out((nt, 16, 3), npdt)
    check(lib().lc_rho_row(prec, _p(row, row.shape, row.dtype), _p(mom5, (nt, 5, nv, nv), row.dtype),
                           tf, nt, nv, po))
    return out


def rhopi_row(row, diag, p1, p2, p12x5, tf, idir):
    nt, nv = _dims(row)
    prec, npdt = _prec(row)
    m = (nt, nv, nv)
    out, po = _out((nt, 4), npdt)
    check(lib().lc_rhopi_row(prec, _p(row, row.shape, row.dtype), _p(diag, row.shape, row.dtype),
                             _p(p1, m, row.dtype), _p(p2, m, row.dtype),
                             _p(p12x5, (nt, 5, nv, nv), row.dtype), tf, nt, nv, idir, po))
    return out


def pipi_row(iso, row, diag, m1, m2, m3, m4, tf):
    nt, nv = _dims(row)
    prec, npdt = _prec(row)
    m = (nt, nv, nv)
    out, po = _out((nt, 5), npdt)
    check(lib().lc_pipi_row(prec, iso, _p(row, row.shape, row.dtype), _p(diag, row.shape, row.dtype),
                            _p(m1, m, row.dtype), _p(m2, m, row.dtype), _p(m3, m, row.dtype),
                            _p(m4, m, row.dtype), tf, nt, nv, po))
    return out


def loop_unity(diag, mom):
    nt, nv = _dims(diag)
    prec, npdt = _prec(diag)
    out, po = _out((nt,), npdt)
    check(lib().lc_loop_unity(prec, _p(diag, diag.shape, diag.dtype), _p(mom, (nt, nv, nv), diag.dtype),
                              nt, nv, po))
    return out


def loop_unity_noalloc(diag, mom, out):
    """lc_loop_unity_noalloc: the loop_unity trace into the DeviceArray out [nt] with no allocation
    (partial sums in the library's fixed device array); ordered on the library stream, not
    synchronised."""
    nt, nv = _dims(diag)
    prec, npdt = _prec(diag)
    check(lib().lc_loop_unity_noalloc(prec, _p(diag, diag.shape, diag.dtype), _p(mom, (nt, nv, nv), diag.dtype),
                                      nt, nv, _p(out, (nt,), diag.dtype)))
    return out


def sigma_row(row, x2, tf):
    nt, nv = _dims(row)
    prec, npdt = _prec(row)
    out, po = _out((nt, 4), npdt)
    check(lib().lc_sigma_row(prec, _p(row, row.shape, row.dtype), _p(x2, (nt, 2, nv, nv), row.dtype),
                             tf, nt, nv, po))
    return out


def sigmapi_fish_hor(diag, p1, p2):
    nt, nv = _dims(diag)
    prec, npdt = _prec(diag)
    m = (nt, nv, nv)
    out, po = _out((nt,), npdt)
    check(lib().lc_sigmapi_fish_hor(prec, _p(diag, diag.shape, diag.dtype), _p(p1, m, diag.dtype),
                                    _p(p2, m, diag.dtype), nt, nv, po))
    return out


def sigmapi_triangle_row(row, diag, p1, p2, x2, tf):
    nt, nv = _dims(row)
    prec, npdt = _prec(row)
    m = (nt, nv, nv)
    out, po = _out((nt, 2), npdt)
    check(lib().lc_sigmapi_triangle_row(prec, _p(row, row.shape, row.dtype), _p(diag, row.shape, row.dtype),
                                        _p(p1, m, row.dtype), _p(p2, m, row.dtype),
                                        _p(x2, (nt, 2, nv, nv), row.dtype), tf, nt, nv, po))
    return out


def selftest_mfma(a, b, prec=DOUBLE):
    """A (16x4) * B (4x16) on one matrix-core instruction. Returns (D as the kernels map the
    accumulators, raw accumulators [64 lanes][4])."""
    dt = np.float64 if prec == DOUBLE else np.float32
    a = np.ascontiguousarray(a, dt)
    b = np.ascontiguousarray(b, dt)
    d = np.zeros(512, dt)
    check(lib().lc_selftest_mfma(prec, a.ctypes.data, b.ctypes.data, d.ctypes.data))
    return d[:256].reshape(16, 16), d[256:].reshape(64, 4)
