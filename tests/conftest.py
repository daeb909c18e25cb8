import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU case")


@pytest.fixture
def lc():
    """The HIP library, initialised on device 0 (re-initialised if a test finalised it)."""
    from contractions_code_scattering_amd import rows
    rows.init(0)
    yield rows


def relerr(got, ref):
    """max |got - ref| / max |ref| (normwise over the array)."""
    import numpy as np
    got = np.asarray(got)
    ref = np.asarray(ref)
    den = np.abs(ref).max()
    return float(np.abs(got - ref).max() / den) if den > 0 else float(np.abs(got).max())


def elemerr(got, ref, floor=1e-3):
    """max over elements of |got-ref| / max(|ref|, floor * max|ref|)."""
    import numpy as np
    got = np.asarray(got)
    ref = np.asarray(ref)
    den = np.maximum(np.abs(ref), floor * np.abs(ref).max())
    return float((np.abs(got - ref) / den).max())
