"""GPU covariant derivative and the drivers' five elemental types (lc_covariant_derivative,
lc_build_elementals_types) against the oracle's restatement (tests/compute_rho.cpp:116-229).

FP64 bars: derivative 1e-13, types 1e-12 relative per type. Also: type 0 is bit-identical to
lc_build_elementals, and the GPU types are gauge invariant (random SU(3) gauge transformation).
"""
import numpy as np
import pytest

import oracle as O
from conftest import relerr
from test_elemental_types import _lattice, _neighbour, _su3

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nt,nv,nx,ny,nz,np_", [(2, 3, 3, 4, 5, 2), (1, 70, 8, 8, 6, 1), (2, 8, 4, 4, 4, 3)])
def test_types_vs_oracle(lc, nt, nv, nx, ny, nz, np_):
    from contractions_code_scattering_amd import engine
    from contractions_code_scattering_amd.devmem import to_device
    U, V, ph = _lattice(10 + nv, nt, nv, nx, ny, nz, np_)
    print("elemental types nt=%d nv=%d lattice %dx%dx%d np=%d" % (nt, nv, nx, ny, nz, np_), flush=True)
    dU, dV_, dph = to_device(U), to_device(V), to_device(ph)
    dv = engine.covariant_derivative(dU, dV_, nx, ny, nz)
    ref_dv = O.covariant_derivative(U, V, nx, ny, nz)
    got_dv = dv.to_numpy()
    for j in range(3):
        assert relerr(got_dv[j], ref_dv[j]) < 1e-13, (j, relerr(got_dv[j], ref_dv[j]))
    T = engine.build_elementals_types(dV_, dv, dph).to_numpy()
    R = O.elemental_types(V, ref_dv, ph)
    for k in range(5):
        assert relerr(T[:, k], R[:, k]) < 1e-12, (k, relerr(T[:, k], R[:, k]))
    plain = engine.build_elementals(dV_, dph).to_numpy()
    assert np.array_equal(T[:, 0], plain)


def test_types_gauge_invariant(lc):
    from contractions_code_scattering_amd import engine
    from contractions_code_scattering_amd.devmem import to_device
    nt, nv, nx, ny, nz = 2, 6, 4, 3, 5
    U, V, ph = _lattice(20, nt, nv, nx, ny, nz, 2)
    ns = nx * ny * nz
    G = _su3(np.random.default_rng(21), nt * ns).reshape(nt, ns, 3, 3)
    Ug = np.empty_like(U)
    for j in range(3):
        Ug[:, j] = G @ U[:, j] @ np.conj(np.swapaxes(G[:, _neighbour(nx, ny, nz, j, +1)], -1, -2))
    Vg = np.einsum("tsab,tcsb->tcsa", G, V.reshape(nt, nv, ns, 3)).reshape(nt, nv, 3 * ns)
    out = []
    for u, v in ((U, V), (Ug, np.ascontiguousarray(Vg))):
        dv = to_device(v)
        out.append(engine.build_elementals_types(dv, engine.covariant_derivative(to_device(u), dv, nx, ny, nz),
                                                 to_device(ph)).to_numpy())
    for k in range(5):
        assert relerr(out[1][:, k], out[0][:, k]) < 1e-12, k


def test_types_reject_bad_shapes(lc):
    from contractions_code_scattering_amd import engine
    from contractions_code_scattering_amd.devmem import to_device
    U, V, ph = _lattice(30, 1, 2, 2, 2, 2, 1)
    with pytest.raises(ValueError, match="links"):
        engine.covariant_derivative(to_device(U[:, :2].copy()), to_device(V), 2, 2, 2)
    with pytest.raises(ValueError, match="evec"):
        engine.covariant_derivative(to_device(U), to_device(V), 2, 2, 3)


def test_types_gauge_invariant_full_lattice(lc):
    """At the benchmark's lattice (L = 24, Nv = 64; Lt = 2 to keep the host transformation short):
    the five types of a random SU(3) gauge transform of links and eigenvectors equal those of the
    untransformed ones, and zero momentum keeps type 1 Hermitian and types 2-4 anti-Hermitian."""
    from contractions_code_scattering_amd import engine
    from contractions_code_scattering_amd.devmem import to_device
    nt, nv, L = 2, 64, 24
    U, V, ph = _lattice(40, nt, nv, L, L, L, 2)
    ph[1] = 1.0  # second momentum: p = 0
    ns = L ** 3
    G = _su3(np.random.default_rng(41), nt * ns).reshape(nt, ns, 3, 3)
    Ug = np.empty_like(U)
    for j in range(3):
        Ug[:, j] = G @ U[:, j] @ np.conj(np.swapaxes(G[:, _neighbour(L, L, L, j, +1)], -1, -2))
    Vg = np.ascontiguousarray(np.einsum("tsab,tcsb->tcsa", G, V.reshape(nt, nv, ns, 3)).reshape(nt, nv, 3 * ns))
    print("elemental types gauge invariance L=%d nv=%d" % (L, nv), flush=True)
    out = []
    for u, v in ((U, V), (Ug, Vg)):
        dv = to_device(v)
        out.append(engine.build_elementals_types(dv, engine.covariant_derivative(to_device(u), dv, L, L, L),
                                                 to_device(ph)).to_numpy())
    for k in range(5):
        assert relerr(out[1][:, k], out[0][:, k]) < 1e-12, (k, relerr(out[1][:, k], out[0][:, k]))
    T = out[0][1]  # p = 0
    for t in range(nt):
        assert relerr(T[1, t], np.conj(T[1, t].T)) < 1e-12
        for j in range(3):
            assert relerr(T[2 + j, t], -np.conj(T[2 + j, t].T)) < 1e-12


@pytest.mark.parametrize("c", [1.0, 0.25, 0.3])
def test_normalisation_scaling_laws(lc, c):
    """lc_covariant_derivative_scaled: dV(c) = (c / (1/2)) dV(1/2), bit-exact for power-of-two ratios
    and against the oracle with the same c; type 0 independent of c, type 1 ~ c^2, types 2-4 ~ c."""
    from contractions_code_scattering_amd import engine
    from contractions_code_scattering_amd.devmem import to_device
    nt, nv, nx, ny, nz = 2, 8, 4, 4, 4
    U, V, ph = _lattice(40, nt, nv, nx, ny, nz, 2)
    dU, dV_, dph = to_device(U), to_device(V), to_device(ph)
    print("covariant derivative normalisation c=%g" % c, flush=True)
    d5 = engine.covariant_derivative(dU, dV_, nx, ny, nz)
    dc = engine.covariant_derivative(dU, dV_, nx, ny, nz, c=c)
    a5, ac = d5.to_numpy(), dc.to_numpy()
    r = c / 0.5
    if r in (2.0, 0.5):
        assert np.array_equal(ac, r * a5)
    assert relerr(ac, O.covariant_derivative(U, V, nx, ny, nz, c=c)) < 1e-13
    T5 = engine.build_elementals_types(dV_, d5, dph).to_numpy()
    Tc = engine.build_elementals_types(dV_, dc, dph).to_numpy()
    assert np.array_equal(Tc[:, 0], T5[:, 0])
    if r in (2.0, 0.5):
        assert np.array_equal(Tc[:, 1], r * r * T5[:, 1])
        assert np.array_equal(Tc[:, 2:], r * T5[:, 2:])
    else:
        assert relerr(Tc[:, 1], r * r * T5[:, 1]) < 1e-13
        assert relerr(Tc[:, 2:], r * T5[:, 2:]) < 1e-13
