"""GPU perambulator reader (lc_read_prop_device) against the oracle's restatement of readPropH
(lib/contract.cpp:91-160): exact for FP64, FP32 rounding for PRECISION SINGLE; error paths."""
import numpy as np
import pytest

import oracle as O
from contractions_code_scattering_amd import inputs as I

pytestmark = pytest.mark.gpu


def _write(path, nt, nv, seed, cfg=b"cfg-test"):
    g = I.rng(seed)
    body = I.cn(g, (4, nv, nt, 4, nv, nt))  # file order [s1][v1][t1][s2][v2][t2]
    with open(path, "wb") as f:
        f.write(cfg.ljust(100, b"\0"))
        f.write(np.array([nt, nv], dtype="<u4").tobytes())
        f.write(body.astype("<c16").tobytes())
    return body


@pytest.mark.parametrize("nt,nv", [(3, 4), (4, 8), (2, 25)])
def test_read_prop_device_matches_oracle(lc, tmp_path, nt, nv):
    from contractions_code_scattering_amd.devmem import read_perambulator
    path = tmp_path / "prop.bin"
    _write(path, nt, nv, 7 + nv)
    ref = O.read_prop_file(str(path), nt, nv)
    got = read_perambulator(path, nt, nv).to_numpy()
    np.testing.assert_array_equal(got, ref)
    got32 = read_perambulator(path, nt, nv, np.complex64).to_numpy()
    np.testing.assert_array_equal(got32, ref.astype(np.complex64))


def test_read_prop_device_errors(lc, tmp_path):
    from contractions_code_scattering_amd import LcError
    from contractions_code_scattering_amd.devmem import read_perambulator
    path = tmp_path / "prop.bin"
    _write(path, 3, 4, 1)
    with pytest.raises(LcError) as e:
        read_perambulator(path, 3, 5)
    assert e.value.code == -7
    with pytest.raises(LcError) as e:
        read_perambulator(tmp_path / "missing.bin", 3, 4)
    assert e.value.code == -7
    short = tmp_path / "short.bin"
    short.write_bytes(path.read_bytes()[:-16])
    with pytest.raises(LcError) as e:
        read_perambulator(short, 3, 4)
    assert e.value.code == -7
    # the library stays usable after a failed read
    np.testing.assert_array_equal(read_perambulator(path, 3, 4).to_numpy(), O.read_prop_file(str(path), 3, 4))
