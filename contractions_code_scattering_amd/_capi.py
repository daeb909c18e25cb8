"""ctypes binding of libcontract_hip.so (the C-ABI declared in include/lapH_contract.h).

The library must have been built in-tree (``python -m contractions_code_scattering_amd.build``);
there is no fallback: loading fails loudly when it is missing.
"""
import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcontract_hip.so")
HEADER = os.path.join(HERE, "..", "include", "lapH_contract.h")

SINGLE, DOUBLE = 0, 1
I0, I1 = 0, 1
ERRORS = {0: "LC_OK", -1: "LC_ERR_NOT_INIT", -2: "LC_ERR_ARG", -3: "LC_ERR_HIP",
          -4: "LC_ERR_ALREADY", -5: "LC_ERR_NODEV", -6: "LC_ERR_NOMEM", -7: "LC_ERR_IO"}


class LcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


_lib = None

_VP, _I, _D, _LL = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_longlong
_SIGS = {
    "lc_init": [_I], "lc_finalize": [], "lc_is_init": [], "lc_device_count": [],
    "lc_rho_row": [_I, _VP, _VP, _I, _I, _I, _VP],
    "lc_rhopi_row": [_I, _VP, _VP, _VP, _VP, _VP, _I, _I, _I, _I, _VP],
    "lc_pipi_row": [_I, _I, _VP, _VP, _VP, _VP, _VP, _VP, _I, _I, _I, _VP],
    "lc_loop_unity": [_I, _VP, _VP, _I, _I, _VP],
    "lc_sigma_row": [_I, _VP, _VP, _I, _I, _I, _VP],
    "lc_sigmapi_fish_hor": [_I, _VP, _VP, _VP, _I, _I, _VP],
    "lc_sigmapi_triangle_row": [_I, _VP, _VP, _VP, _VP, _VP, _I, _I, _I, _VP],
    "lc_selftest_mfma": [_I, _VP, _VP, _VP],
    "lc_contract_init": [_I, _I, _I], "lc_contract_destroy": [], "lc_set_binary_output": [_I],
    "lc_set_i_dir": [_I], "lc_read_prop": [ctypes.c_char_p, _I, ctypes.POINTER(_VP)],
    "lc_calculate_rho": [_I, _I, _VP, _VP, _I, ctypes.c_char_p],
    "lc_calculate_rhopi": [_I, _I, _I, _VP, _VP, _VP, _VP, _I, ctypes.c_char_p],
    "lc_calculate_pipi": [_I, _I, _VP, _VP, _VP, _VP, _VP, _I, ctypes.c_char_p, _I],
    "lc_calculate_sigma": [_I, _I, _VP, _VP, _I, ctypes.c_char_p],
    "lc_calculate_sigmapi": [_I, _I, _VP, _VP, _VP, _VP, _I, ctypes.c_char_p],
}


def lib():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libcontract_hip.so is not built: run "
                              "`python -m contractions_code_scattering_amd.build` (no fallback path)")
        L = ctypes.CDLL(LIB_PATH)
        L.lc_last_error.restype = ctypes.c_char_p
        L.lc_last_error.argtypes = []
        L.lc_free_host.restype = None
        L.lc_free_host.argtypes = [_VP]
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _I
        for name, (args, res) in _EXTRA_SIGS.items():
            if hasattr(L, name):
                f = getattr(L, name)
                f.argtypes = args
                f.restype = res
        _lib = L
    return _lib


_EXTRA_SIGS = {}


def register(name, args, restype=ctypes.c_int):
    _EXTRA_SIGS[name] = (args, restype)
    if _lib is not None and hasattr(_lib, name):
        f = getattr(_lib, name)
        f.argtypes = args
        f.restype = restype


def check(code):
    if code != 0:
        raise LcError(code, lib().lc_last_error().decode(errors="replace"))
    return code


def declared_functions(header=HEADER):
    """Names of the functions declared in include/lapH_contract.h."""
    txt = open(header).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(lc_[a-z0-9_]+)\s*\(", txt)))
