"""Golden fixtures (tests/golden/rows_nv*_nt*.npz): seeded inputs and the oracle's FP64 outputs of
every sink-row diagram at Nv in {4, 8}, Nt = 4.

CPU: the oracle re-run on the stored inputs reproduces the stored outputs (the fixtures and the
oracle stay in step), and the stored inputs are the seeded generator's.
GPU: the HIP rows through the C-ABI match the stored outputs to 1e-12 (normwise per array).
"""
import glob
import os

import numpy as np
import pytest

from conftest import relerr

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "rows_nv*_nt*.npz")))
TOL = 1e-12


def _load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_fixtures_present():
    assert len(FIXTURES) >= 2


@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_oracle_reproduces_golden(path):
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_golden as M
    f = _load(path)
    nv, nt = f["tau"].shape[-1], f["tau"].shape[0]
    regen = M.make_inputs(nv, nt, int(f["seed"]))
    for k in ("tau", "mats", "mom5", "x2"):
        np.testing.assert_array_equal(regen[k], f[k])
    out = M.oracle_outputs(f)
    for k, v in out.items():
        assert relerr(v, f["out_" + k]) < 1e-14, k


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_gpu_rows_match_golden(lc, path):
    from contractions_code_scattering_amd.devmem import DeviceArray
    f = _load(path)
    tau, (p1, p2, p3, p4), mom5, x2 = f["tau"], f["mats"], f["mom5"], f["x2"]
    nt = tau.shape[0]
    diag = np.ascontiguousarray(tau[np.arange(nt), np.arange(nt)])
    d = DeviceArray.from_numpy
    dd, d1, d2, d3, d4, dm5, dx2 = d(diag), d(p1), d(p2), d(p3), d(p4), d(mom5), d(x2)
    got = dict(rho=[], rhopi=[[], [], []], pipi=[[], []], sigma=[], sigmapi_triangle=[])
    for tf in range(nt):
        row = d(tau[tf])
        print("golden %s tf=%d" % (os.path.basename(path), tf), flush=True)
        got["rho"].append(lc.rho_row(row, dm5, tf))
        for idir in range(3):
            got["rhopi"][idir].append(lc.rhopi_row(row, dd, d1, d2, dm5, tf, idir))
        for j, iso in enumerate((1, 0)):
            got["pipi"][j].append(lc.pipi_row(iso, row, dd, d1, d2, d3, d4, tf))
        got["sigma"].append(lc.sigma_row(row, dx2, tf))
        got["sigmapi_triangle"].append(lc.sigmapi_triangle_row(row, dd, d1, d2, dx2, tf))
    got = {k: np.array(v) for k, v in got.items()}
    got["loop_unity"] = lc.loop_unity(dd, d1)
    got["sigmapi_fish_hor"] = lc.sigmapi_fish_hor(dd, d1, d2)
    for k, v in got.items():
        err = relerr(v, f["out_" + k])
        assert err <= TOL, (k, err)
