"""GPU parity of the batched engine against the oracle's row-by-row restatement.

rho_i is the (l,k) = (0,0) op pair of run_ContractRho with Mom type 0 = Phi(p); the pi-pi
traces are those of run_ContractPiPi_I0 with P1 = P4 = Phi(p), P2 = P3 = Phi(p)^dag.
FP64 bar: 1e-12 relative per output array.
"""
import numpy as np
import pytest

import oracle as O
from contractions_code_scattering_amd import inputs as I
from conftest import elemerr, relerr

pytestmark = pytest.mark.gpu


def _inputs(seed, nt, nv, np_):
    g = I.rng(seed)
    tau = I.perambulator(g, nt, nv, 0.3)
    phi = np.stack([I.random_elemental(g, nt, nv) for _ in range(np_)])
    return tau, phi


def _oracle(tau, phi, p, src):
    """[nsrc][nt][8] correlators of momentum p from the oracle (weights removed)."""
    nt, nv = tau.shape[0], tau.shape[-1]
    P = phi[p]
    Pd = I.dagger(P)
    mom5 = np.ascontiguousarray(np.stack([P] * 5, axis=1))
    diag = O.diag_of(tau)
    out = np.zeros((len(src), nt, 8), complex)
    for tf in range(nt):
        r = O.rho_row(tau[tf], mom5, tf)            # [ti][16][3]
        q = O.pipi_row(0, tau[tf], diag, P, Pd, Pd, P, tf)  # I0 weights
        for k, t0 in enumerate(src):
            out[k, tf, 0:3] = r[t0, 0, :]
            out[k, tf, 3] = q[t0, 0] * -3.0
            out[k, tf, 4] = q[t0, 1] * -3.0
            out[k, tf, 5] = q[t0, 4] * (-3.0 / 5.0)
            out[k, tf, 6] = q[t0, 2]
            out[k, tf, 7] = q[t0, 3]
    return out


@pytest.mark.parametrize("nt,nv,np_,src0,step", [(4, 4, 2, 0, 1), (5, 8, 1, 1, 2), (3, 16, 2, 0, 1),
                                                  (4, 25, 1, 0, 2), (2, 128, 1, 0, 1)])
def test_engine_vs_oracle(lc, nt, nv, np_, src0, step):
    from contractions_code_scattering_amd import engine
    from contractions_code_scattering_amd.devmem import to_device
    tau, phi = _inputs(11, nt, nv, np_)
    src = list(range(src0, nt, step))
    raw, corr, avg = engine.rho_pipi(to_device(tau), to_device(phi), src0=src0, src_step=step)
    corr = corr.to_numpy()
    avg = avg.to_numpy()
    for p in range(np_):
        ref = _oracle(tau, phi, p, src)
        for d in range(8):
            assert relerr(corr[p, :, :, d], ref[:, :, d]) < 1e-12, (p, d)
            assert elemerr(corr[p, :, :, d], ref[:, :, d]) < 1e-12, (p, d)
        # device source average vs host restatement of dumpData's average over the source list
        for d in range(8):
            c = np.array([np.mean([ref[k, (t0 + dt) % nt, d] for k, t0 in enumerate(src)])
                          for dt in range(nt)])
            assert relerr(avg[p, d], c) < 1e-12


def test_engine_chunking_is_invisible(lc):
    """A scratch budget that forces one momentum per chunk gives identical results."""
    from contractions_code_scattering_amd import engine
    from contractions_code_scattering_amd.devmem import to_device
    nt, nv, np_ = 3, 8, 3
    tau, phi = _inputs(12, nt, nv, np_)
    t, f = to_device(tau), to_device(phi)
    _, c1, _ = engine.rho_pipi(t, f)
    per_p = (6 * nt * nt + 2 * nt + 4 * nt) * 16 * nv * nv * 16
    _, c2, _ = engine.rho_pipi(t, f, max_scratch_bytes=per_p)
    assert np.array_equal(c1.to_numpy(), c2.to_numpy())


def _bf16_exact(g, shape):
    import numpy as np
    x = g.standard_normal(shape).astype(np.float32)
    u = x.view(np.uint32) & np.uint32(0xFFFF0000)
    return u.view(np.float32)


def test_mfma_bf16_layout(lc):
    from contractions_code_scattering_amd import engine
    g = I.rng(21)
    a = _bf16_exact(g, (16, 32))
    b = _bf16_exact(g, (32, 16))
    ref = a.astype(np.float64) @ b.astype(np.float64)
    d, raw = engine.selftest_mfma_bf16(a, b)
    print("raw accumulator rows for lanes 0,16,32,48:",
          [[int(np.argmin(np.abs(ref[:, l & 15] - raw[l, q]))) for q in range(4)] for l in (0, 16, 32, 48)])
    np.testing.assert_allclose(d, ref, rtol=1e-6, atol=1e-5)


@pytest.mark.parametrize("form", ["exact", "bf16x3"])
@pytest.mark.parametrize("nt,nv,np_,src0,step", [(4, 16, 2, 0, 1), (3, 64, 1, 0, 1), (4, 64, 3, 0, 1),
                                                  (5, 64, 2, 1, 2), (4, 33, 1, 0, 1), (2, 128, 1, 0, 1)])
def test_engine_mixed_vs_fp64(lc, form, nt, nv, np_, src0, step):
    """Both mixed forms: the exact one within 1e-6 per correlator array, BF16x3 (at Nv = 64 the
    fused-sandwich chain, other Nv the four-level chain) within 2e-6."""
    from contractions_code_scattering_amd import engine
    from contractions_code_scattering_amd.devmem import to_device
    bar = 1e-6 if form == "exact" else 2e-6
    tau, phi = _inputs(13, nt, nv, np_)
    t, f = to_device(tau), to_device(phi)
    _, c64, a64 = engine.rho_pipi(t, f, src0=src0, src_step=step)
    engine.set_mixed_form(engine.MIXED_EXACT if form == "exact" else engine.MIXED_BF16X3)
    try:
        _, cmx, amx = engine.rho_pipi(t, f, mixed=True, src0=src0, src_step=step)
    finally:
        engine.set_mixed_form(engine.MIXED_EXACT)
    c64, cmx = c64.to_numpy(), cmx.to_numpy()
    errs = engine.rel_errors(amx.to_numpy(), a64.to_numpy())
    print("mixed (%s) vs fp64:" % form, errs)
    for d in range(8):
        assert relerr(cmx[..., d], c64[..., d]) < bar, (d, relerr(cmx[..., d], c64[..., d]))
    assert max(e["max_normwise"] for e in errs.values()) < bar
