"""GPU parity of the whole-call drop-ins (initContract + calculate* through the C wrappers)
against the oracle's restatement of the reference flows (lib/interface.cpp + lib/contract.cpp).

The FP64 side channel carries the printed values at full precision (bar 1e-12 relative per
correlator block); the %+e text must match the oracle's text line structure exactly and its
numbers to the 7 printed digits.
"""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
from contractions_code_scattering_amd import _capi, check
from contractions_code_scattering_amd import inputs as I
from conftest import relerr

pytestmark = pytest.mark.gpu
NT, NV = 4, 8


@pytest.fixture
def dropin(lc):
    L = _capi.lib()
    check(L.lc_contract_init(NV, NT, 0))
    check(L.lc_set_binary_output(1))
    yield L
    check(L.lc_set_binary_output(0))
    check(L.lc_contract_destroy())


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _compare(path, text_ref, vals_ref):
    text = open(path).read()
    got = np.fromfile(path + ".bin", dtype="<f8")
    ref = np.concatenate([np.asarray(v).view(np.float64).ravel() for v in vals_ref])
    assert got.shape == ref.shape
    assert relerr(got, ref) < 1e-12
    num = re.compile(r"[+-]\d\.\d{6}e[+-]\d+")
    assert num.sub("#", text) == num.sub("#", text_ref)
    a = np.array([float(x) for x in num.findall(text)])
    b = np.array([float(x) for x in num.findall(text_ref)])
    assert np.all(np.abs(a - b) <= 2e-6 * np.abs(b) + 1e-300 + 1e-6 * np.abs(b).max())


def _tau(seed):
    return I.perambulator(I.rng(seed), NT, NV, 0.4)


def test_calculate_rho(dropin, tmp_path):
    tau = _tau(1)
    mom = np.stack([I.hermitian_elemental(I.rng(2 + k), NT, NV) for k in range(5)])
    out = str(tmp_path / "rho.dat")
    check(dropin.lc_calculate_rho(NV, NT, _p(tau), _p(mom), 1, out.encode()))
    _, corr, text = O.calculate_rho(tau, mom)
    vals = [np.ascontiguousarray(corr.transpose(1, 0, 2))]  # printed iop-major, then it, then dir
    _compare(out, text, vals)


def test_calculate_rhopi(dropin, tmp_path):
    tau = _tau(3)
    g = I.rng(4)
    p1, p2 = I.random_elemental(g, NT, NV), I.random_elemental(g, NT, NV)
    p12 = np.stack([I.random_elemental(g, NT, NV) for _ in range(5)])
    out = str(tmp_path / "rhopi.dat")
    check(dropin.lc_set_i_dir(1))
    check(dropin.lc_calculate_rhopi(1, NV, NT, _p(tau), _p(p1), _p(p2), _p(p12), 1, out.encode()))
    _, corr, text = O.calculate_rhopi(1, tau, p1, p2, p12)
    _compare(out, text, [2 * np.ascontiguousarray(corr.T)])


@pytest.mark.parametrize("iso", [0, 1])
def test_calculate_pipi(dropin, tmp_path, iso):
    tau = _tau(5)
    g = I.rng(6)
    ms = [I.random_elemental(g, NT, NV) for _ in range(4)]
    out = str(tmp_path / "pipi.dat")
    check(dropin.lc_calculate_pipi(NV, NT, _p(tau), *[_p(m) for m in ms], 1, out.encode(), iso))
    _, corr, text = O.calculate_pipi(tau, *ms, iso)
    nd = 5 if iso == 0 else 4
    _compare(out, text, [np.ascontiguousarray(corr[:, :nd].T)])


def test_calculate_sigma(dropin, tmp_path):
    tau = _tau(7)
    g = I.rng(8)
    mom = np.stack([I.random_elemental(g, NT, NV) for _ in range(4)])
    out = str(tmp_path / "sigma.dat")
    check(dropin.lc_calculate_sigma(NV, NT, _p(tau), _p(mom), 1, out.encode()))
    loops, _, corr, text = O.calculate_sigma(tau, mom)
    _compare(out, text, list(loops) + [-np.ascontiguousarray(corr.T)])


def test_calculate_sigmapi(dropin, tmp_path):
    tau = _tau(9)
    g = I.rng(10)
    p1, p2 = I.random_elemental(g, NT, NV), I.random_elemental(g, NT, NV)
    p12 = np.stack([I.random_elemental(g, NT, NV) for _ in range(2)])
    out = str(tmp_path / "sigmapi.dat")
    check(dropin.lc_calculate_sigmapi(NV, NT, _p(tau), _p(p1), _p(p2), _p(p12), 1, out.encode()))
    loops, fh, _, corr, text = O.calculate_sigmapi(tau, p1, p2, p12)
    f = -2.0 / np.sqrt(6.0)
    _compare(out, text, list(loops) + [fh, f * np.ascontiguousarray(corr.T)])


def test_single_precision_rho(dropin, tmp_path):
    tau = _tau(11)
    mom = np.stack([I.hermitian_elemental(I.rng(12 + k), NT, NV) for k in range(5)])
    out = str(tmp_path / "rho32.dat")
    tau32 = tau.astype(np.complex64)
    check(dropin.lc_calculate_rho(NV, NT, _p(tau32), _p(mom), 0, out.encode()))
    _, corr, _ = O.calculate_rho(tau, mom)
    got = np.fromfile(out + ".bin", dtype="<f8")
    ref = np.ascontiguousarray(corr.transpose(1, 0, 2)).view(np.float64).ravel()
    assert relerr(got, ref) < 1e-5


def test_read_prop_matches_oracle(dropin, tmp_path):
    g = I.rng(13)
    body = I.cn(g, (4, NV, NT, 4, NV, NT))
    path = tmp_path / "prop.bin"
    with open(path, "wb") as f:
        f.write(b"cfg-42".ljust(100, b"\0"))
        f.write(np.array([NT, NV], "<u4").tobytes())
        f.write(body.astype("<c16").tobytes())
    ptr = ctypes.c_void_p()
    check(dropin.lc_read_prop(str(path).encode(), 1, ctypes.byref(ptr)))
    n = NT * NT * 16 * NV * NV
    got = np.ctypeslib.as_array((ctypes.c_double * (2 * n)).from_address(ptr.value)).copy()
    dropin.lc_free_host(ptr)
    ref = O.read_prop_file(str(path), NT, NV)
    assert np.array_equal(got.view(np.complex128).reshape(ref.shape), ref)
    # header mismatch -> LC_ERR_IO, no exit
    bad = tmp_path / "bad.bin"
    with open(bad, "wb") as f:
        f.write(b"x".ljust(100, b"\0"))
        f.write(np.array([NT + 1, NV], "<u4").tobytes())
    assert dropin.lc_read_prop(str(bad).encode(), 1, ctypes.byref(ptr)) == -7
