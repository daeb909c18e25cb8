"""MI355X-native LapH/distillation meson contraction engine.

The compute path is libcontract_hip.so: hand-written HIP kernels for gfx950 behind the C-ABI of
include/lapH_contract.h, with a C++ drop-in of the reference's contract.h / interface.h API.
This package holds the Python-side ctypes binding (used by the tests and the benchmark) and
the seeded synthetic-input generators.
"""
from ._capi import DOUBLE, I0, I1, SINGLE, LcError, check, lib  # noqa: F401
