"""doubleTriangleHor with tau^dag (lc_set_dthor_dagger) against the dense brute force.

The reference's kernel multiplies by conj(tau_ba) element-wise, without transposing the
eigenvector indices (lib/calculate_pipi_doubleTriangleHor_core.h:126). The option replaces that by
the conjugate transpose. The brute force builds explicit 4Nv x 4Nv matrices, T = tau(tf, ti):

    Tr[K(m3(tf)) T K(m1(ti)) T^dag K(m4(tf)) T K(m2(ti)) T^dag],   K(m) = 1_4 (x) m

It is independent of the view machinery the library uses. With the option off, the reference's
value (the brute force with W_ab = conj(tau_ba)) must come back. FP64 bar 1e-12; mixed 2e-6
normwise against the FP64 engine.
"""
import numpy as np
import pytest

import oracle as O
from contractions_code_scattering_amd import inputs as I
from conftest import relerr

pytestmark = pytest.mark.gpu


def _hor(tau, m1, m2, m3, m4, tf, ti, dagger):
    K = O.dense.kron
    T = O.dense.tau(tau[tf, ti])
    Th = T.conj().T
    W = Th if dagger else O.dense.w_hor(tau[tf, ti])
    return np.trace(K(m3[tf]) @ T @ K(m1[ti]) @ W @ K(m4[tf]) @ T @ K(m2[ti]) @ Th)


@pytest.fixture
def dthor(lc):
    from contractions_code_scattering_amd import engine
    engine.set_dthor_dagger(False)
    yield engine
    engine.set_dthor_dagger(False)


def _dev(x):
    from contractions_code_scattering_amd.devmem import DeviceArray
    return DeviceArray.from_numpy(x)


@pytest.mark.parametrize("nt,nv", [(3, 8), (2, 25)])
def test_row_launcher(dthor, nt, nv):
    from contractions_code_scattering_amd import rows
    g = I.rng(41)
    tau = I.perambulator(g, nt, nv, 0.3)
    ms = [I.random_elemental(g, nt, nv) for _ in range(4)]
    diag = O.diag_of(tau)
    for on in (False, True):
        dthor.set_dthor_dagger(on)
        for tf in range(nt):
            got = rows.pipi_row(0, _dev(np.ascontiguousarray(tau[tf])), _dev(diag), *map(_dev, ms), tf)
            ref = np.array([_hor(tau, *ms, tf, ti, on) for ti in range(nt)]) * (-5.0 / 3.0)
            assert relerr(got[:, 4], ref) < 1e-12, (on, tf, relerr(got[:, 4], ref))
            # the other diagrams do not depend on the option
            base = O.pipi_row(0, tau[tf], diag, *ms, tf)
            assert relerr(got[:, :4], base[:, :4]) < 1e-12


@pytest.mark.parametrize("nt,nv,np_", [(3, 16, 2), (3, 64, 2)])
def test_engines(dthor, nt, nv, np_):
    """FP64 engine against the brute force; the mixed engine (fused sandwich at Nv = 64) against
    the FP64 one, with the option on."""
    from contractions_code_scattering_amd.devmem import to_device
    g = I.rng(42)
    tau = I.perambulator(g, nt, nv, 0.3)
    phi = np.stack([I.random_elemental(g, nt, nv) for _ in range(np_)])
    t, f = to_device(tau), to_device(phi)
    _, c_ref, _ = dthor.rho_pipi(t, f)
    dthor.set_dthor_dagger(True)
    _, c64, _ = dthor.rho_pipi(t, f)
    _, cmx, _ = dthor.rho_pipi(t, f, mixed=True)
    c_ref, c64, cmx = c_ref.to_numpy(), c64.to_numpy(), cmx.to_numpy()
    for p in range(np_):
        P = phi[p]
        Pd = I.dagger(P)
        # engine convention: P1 = P4 = Phi(p), P2 = P3 = Phi(p)^dag; corr[p][source][tf][5]
        ref = np.array([[_hor(tau, P, Pd, Pd, P, tf, t0, True) for tf in range(nt)] for t0 in range(nt)])
        assert relerr(c64[p, :, :, 5], ref) < 1e-12, relerr(c64[p, :, :, 5], ref)
        assert relerr(cmx[p, :, :, 5], c64[p, :, :, 5]) < 2e-6
    # every other diagram is unchanged by the option
    keep = [0, 1, 2, 3, 4, 6, 7]
    assert relerr(c64[..., keep], c_ref[..., keep]) < 1e-13
    assert relerr(cmx[..., keep], c_ref[..., keep]) < 2e-6
    # and the option changes dTHor (tau is not symmetric per block)
    assert relerr(c64[..., 5], c_ref[..., 5]) > 1e-3
