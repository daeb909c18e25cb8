"""C-ABI library checks that need no GPU: it loads, exports every function declared in
include/lapH_contract.h, reports the reference's failure modes as error codes, and the drop-in
C++ headers (contract.h / interface.h) compile and link against it for a reference-style driver
that keeps the reference's exit(EXIT_FAILURE) convention."""
import ctypes
import os
import shutil
import subprocess

import pytest

from contractions_code_scattering_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _no_gpu():
    return _capi.lib().lc_device_count() == 0


def test_exports_every_declared_symbol():
    L = _capi.lib()
    names = _capi.declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_not_initialised_errors():
    L = _capi.lib()
    if L.lc_is_init():
        pytest.skip("library already initialised in this process")
    buf = (ctypes.c_double * 16)()
    rc = L.lc_rho_row(1, ctypes.c_void_p(1), ctypes.c_void_p(1), 0, 2, 4, buf)
    assert rc == -1
    assert b"not initialised" in L.lc_last_error()
    assert L.lc_finalize() == -1
    assert L.lc_contract_destroy() == -1


def test_tile_params_and_autotune_arguments():
    """The tile-shape override, the autotuner and graph switches and the dTHor option validate their
    arguments without a GPU."""
    L = _capi.lib()
    for bad in (3, -1, 8):
        assert L.lc_set_tile_params(bad) == -2
        assert b"tiles per workgroup" in L.lc_last_error()
    for ok in (1, 2, 4, 0):
        assert L.lc_set_tile_params(ok) == 0
    assert L.lc_autotune(2) == -2
    assert L.lc_autotune(-1) == 0 and L.lc_autotune(0) == 0
    assert L.lc_autotune_report(None) == -2
    assert L.lc_set_dthor_dagger(2) == -2
    assert L.lc_set_dthor_dagger(1) == 0 and L.lc_set_dthor_dagger(0) == 0
    assert L.lc_engine_graph(2) == -2
    assert L.lc_engine_graph_report(None) == -2
    st = ctypes.c_int(-1)
    assert L.lc_engine_graph_report(ctypes.byref(st)) == 0 and st.value in (0, 1, 2)
    out = (ctypes.c_double * 5)()
    assert L.lc_autotune_report(out) == 0
    assert out[0] in (0.0, 1.0, 2.0, 4.0) and out[1] in (0.0, 1.0, 2.0)


@pytest.mark.skipif(not _no_gpu() if os.path.exists(_capi.LIB_PATH) else True, reason="a GPU is visible")
def test_no_device_errors():
    L = _capi.lib()
    assert L.lc_init(0) == -5
    assert L.lc_contract_init(100, 16, 0) == -5
    assert b"No GPU" in L.lc_last_error() or b"no GPU" in L.lc_last_error()


def test_drop_in_headers_build_a_driver(tmp_path):
    """A reference-style driver (tests/compute_*.cpp shape) compiles against include/ and links."""
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    src = tmp_path / "driver.cpp"
    src.write_text(r'''
#include <interface.h>
int main() {
  contract::contractInfo info;
  info.NeV = 8; info.Nt = 4; info.dev = 0; info.PropH_fileName = "none";
  contract::initContract(info);     // no GPU here: prints the reason and exits with EXIT_FAILURE
  void* prop = contract::readPropH(info.PropH_fileName, contract::DOUBLE);
  calculateRho(info, prop, (double*)0, contract::DOUBLE, "out.dat");
  contract::destroyContract();
  return 0;
}
''')
    exe = tmp_path / "driver"
    libdir = os.path.dirname(_capi.LIB_PATH)
    r = subprocess.run([hipcc, "-std=c++17", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        "-L" + libdir, "-lcontract_hip", "-Wl,-rpath," + libdir],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    if _no_gpu():
        r = subprocess.run([str(exe)], capture_output=True, text=True, cwd=tmp_path, timeout=60)
        assert r.returncode == 1
        assert "No GPU devices found" in r.stderr
