"""lc_loop_unity_noalloc (the allocation-free loop_unity: partial sums in the library's fixed device
array) against lc_loop_unity through the C-ABI: two calls in a row, the second with smaller inputs
(the array keeps the first call's contents, which the second must not see), in both precisions;
and the rejection of sizes beyond the fixed array.

Bar: FP64 1e-13 relative (the two entry points sum in different orders); FP32 inputs 2e-6 (the
allocation-free form accumulates in FP64, the other in the compute type).
"""
import numpy as np
import pytest

from contractions_code_scattering_amd import _capi
from contractions_code_scattering_amd import inputs as I
from conftest import relerr

pytestmark = pytest.mark.gpu


def _case(seed, nt, nv, dtype):
    from contractions_code_scattering_amd.devmem import DeviceArray
    g = I.rng(seed)
    tau = I.perambulator(g, nt, nv, 0.3)
    diag = np.ascontiguousarray(I.diag_of(tau)).astype(dtype)
    mom = I.random_elemental(g, nt, nv).astype(dtype)
    return DeviceArray.from_numpy(diag), DeviceArray.from_numpy(mom), DeviceArray((nt,), dtype)


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-13), (np.complex64, 2e-6)])
def test_loop_unity_noalloc_matches(lc, dtype, tol):
    from contractions_code_scattering_amd.devmem import synchronize
    for seed, nt, nv in ((1, 48, 64), (2, 7, 25)):  # the second call smaller than the first
        print("loop_unity_noalloc nt=%d nv=%d %s" % (nt, nv, np.dtype(dtype).name), flush=True)
        diag, mom, out = _case(seed, nt, nv, dtype)
        ref = lc.loop_unity(diag, mom)
        lc.loop_unity_noalloc(diag, mom, out)
        synchronize()
        got = out.to_numpy()
        assert relerr(got, ref) < tol, (nt, nv, relerr(got, ref))


def test_loop_unity_noalloc_rejects_oversize(lc):
    from contractions_code_scattering_amd.devmem import DeviceArray
    L = _capi.lib()
    d = DeviceArray((1,), np.complex128)
    # 4096 slices of NeV = 1024 need 4096 * 64 partial sums: refused before any launch
    rc = L.lc_loop_unity_noalloc(1, d.ptr, d.ptr, 4096, 1024, d.ptr)
    assert rc == -2, rc
    assert b"above the fixed" in L.lc_last_error()
