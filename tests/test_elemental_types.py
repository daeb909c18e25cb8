"""The oracle's restatement of the drivers' five elemental types (tests/compute_rho.cpp:116-229) and
of the covariant derivative they are built from, pinned by what the mathematics fixes (CPU):

* gauge covariance: under U_j(x) -> g(x) U_j(x) g(x + j)^dag, v -> g v, the derivative transforms
  like v and every type is invariant. A missing dagger, a link taken at the wrong site or a
  neighbour shifted the wrong way breaks this;
* free field (U = 1) on plane waves: nabla_j e^{ik.x} = i sin(k_j) e^{ik.x}, so every type is a
  closed form in the plane waves' momenta;
* zero momentum: the derivative is anti-Hermitian, so type 1 is Hermitian positive semi-definite
  and type 2 + j equals 2 V^dag nabla_j V, anti-Hermitian (a sign error in the commutator gives 0).
"""
import numpy as np

import oracle as O


def _su3(g, n):
    """n random SU(3) matrices [n][3][3]."""
    z = (g.standard_normal((n, 3, 3)) + 1j * g.standard_normal((n, 3, 3))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    q = q * (np.diagonal(r, axis1=1, axis2=2) / np.abs(np.diagonal(r, axis1=1, axis2=2)))[:, None, :]
    return q / np.linalg.det(q)[:, None, None] ** (1.0 / 3.0)


def _lattice(seed, nt, nv, nx, ny, nz, np_):
    g = np.random.default_rng(seed)
    ns = nx * ny * nz
    U = _su3(g, nt * 3 * ns).reshape(nt, 3, ns, 3, 3)
    V = (g.standard_normal((nt, nv, 3 * ns)) + 1j * g.standard_normal((nt, nv, 3 * ns))) / np.sqrt(2)
    ph = np.exp(-1j * g.uniform(0, 2 * np.pi, (np_, ns))).repeat(3, axis=1)
    return U, V, ph


def _coords(nx, ny, nz):
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return x.ravel(), y.ravel(), z.ravel()


def _neighbour(nx, ny, nz, j, step):
    x, y, z = _coords(nx, ny, nz)
    d = [0, 0, 0]
    d[j] = step
    return (x + d[0]) % nx + nx * (((y + d[1]) % ny) + ny * ((z + d[2]) % nz))


def test_gauge_covariance():
    nt, nv, nx, ny, nz = 2, 4, 3, 4, 5
    U, V, ph = _lattice(1, nt, nv, nx, ny, nz, 2)
    ns = nx * ny * nz
    G = _su3(np.random.default_rng(2), nt * ns).reshape(nt, ns, 3, 3)
    Ug = np.empty_like(U)
    for j in range(3):
        fw = _neighbour(nx, ny, nz, j, +1)
        Ug[:, j] = G @ U[:, j] @ np.conj(np.swapaxes(G[:, fw], -1, -2))
    Vg = np.einsum("tsab,tcsb->tcsa", G, V.reshape(nt, nv, ns, 3)).reshape(nt, nv, 3 * ns)
    dV = O.covariant_derivative(U, V, nx, ny, nz)
    dVg = O.covariant_derivative(Ug, Vg, nx, ny, nz)
    # the derivative transforms like the vector
    want = np.einsum("tsab,jtcsb->jtcsa", G, dV.reshape(3, nt, nv, ns, 3)).reshape(dV.shape)
    np.testing.assert_allclose(dVg, want, atol=1e-12)
    np.testing.assert_allclose(O.elemental_types(Vg, dVg, ph), O.elemental_types(V, dV, ph), atol=1e-11)
    # a backward hop with the link at the wrong site, U_j(x) v(x - j), is not covariant (the check
    # has teeth)
    bw = _neighbour(nx, ny, nz, 0, -1)

    def wrong(u, v):
        return np.einsum("tsab,tcsb->tcsa", u[:, 0], v.reshape(nt, nv, ns, 3)[:, :, bw])
    assert np.abs(wrong(Ug, Vg) - np.einsum("tsab,tcsb->tcsa", G, wrong(U, V))).max() > 1e-3


def test_free_field_plane_waves():
    nx, ny, nz = 4, 3, 5
    ns = nx * ny * nz
    x, y, z = _coords(nx, ny, nz)
    # pairs (0, 1), (1, 2) and (3, 4) are joined by p = (1, 0, 3) at equal colour
    ks = [(0, 0, 0), (1, 0, 3), (2, 0, 1), (0, 2, 1), (1, 2, 4), (3, 1, 4)]
    cols = [0, 0, 0, 1, 1, 2]
    V = np.zeros((1, len(ks), ns, 3), complex)
    for c, ((kx, ky, kz), a) in enumerate(zip(ks, cols)):
        V[0, c, :, a] = np.exp(2j * np.pi * (kx * x / nx + ky * y / ny + kz * z / nz))
    V = V.reshape(1, len(ks), 3 * ns)
    U = np.broadcast_to(np.eye(3, dtype=complex), (1, 3, ns, 3, 3)).copy()
    dV = O.covariant_derivative(U, V, nx, ny, nz)
    L = (nx, ny, nz)
    s = np.array([[np.sin(2 * np.pi * k[j] / L[j]) for j in range(3)] for k in ks])  # [c][j]
    for j in range(3):
        np.testing.assert_allclose(dV[j, 0], 1j * s[:, j][:, None] * V[0], atol=1e-12)
    p = (1, 0, 3)
    ph = np.exp(-2j * np.pi * (p[0] * x / nx + p[1] * y / ny + p[2] * z / nz)).repeat(3)[None]
    T = O.elemental_types(V, dV, ph)[0, :, 0]  # [5][r][c]
    # type 0: sum_x e^{i(k_c - k_r - p).x} delta(colour) = ns when k_c = k_r + p (mod L)
    t0 = np.zeros((len(ks), len(ks)), complex)
    for r in range(len(ks)):
        for c in range(len(ks)):
            if cols[r] == cols[c] and all((ks[c][j] - ks[r][j] - p[j]) % L[j] == 0 for j in range(3)):
                t0[r, c] = ns
    assert np.count_nonzero(t0) == 3
    np.testing.assert_allclose(T[0], t0, atol=1e-10)
    np.testing.assert_allclose(T[1], (s @ s.T) * t0, atol=1e-10)
    for j in range(3):
        np.testing.assert_allclose(T[2 + j], 1j * (s[None, :, j] + s[:, None, j]) * t0, atol=1e-10)


def test_zero_momentum_symmetries():
    nt, nv, nx, ny, nz = 2, 5, 3, 3, 4
    U, V, _ = _lattice(3, nt, nv, nx, ny, nz, 1)
    ph = np.ones((1, V.shape[-1]), complex)
    dV = O.covariant_derivative(U, V, nx, ny, nz)
    T = O.elemental_types(V, dV, ph)[0]
    for t in range(nt):
        h = T[1, t]
        np.testing.assert_allclose(h, np.conj(h.T), atol=1e-11)
        assert np.linalg.eigvalsh(h).min() > -1e-10
        for j in range(3):
            c = T[2 + j, t]
            np.testing.assert_allclose(c, 2 * np.conj(V[t]) @ dV[j, t].T, atol=1e-10)
            np.testing.assert_allclose(c, -np.conj(c.T), atol=1e-10)
            assert np.abs(c).max() > 1e-3


def test_normalisation_scaling_laws():
    """The covariant derivative's normalisation c (gwu-qcd's covderiv is not in the reference, so c is
    a parameter): dV scales as c, type 0 does not depend on c, type 1 scales as c^2 and types 2-4 as c;
    power-of-two ratios are exact."""
    nt, nv, nx, ny, nz = 2, 3, 3, 3, 4
    U, V, ph = _lattice(5, nt, nv, nx, ny, nz, 2)
    d5 = O.covariant_derivative(U, V, nx, ny, nz)  # c = 1/2
    for c in (1.0, 0.25, 0.3):
        dc = O.covariant_derivative(U, V, nx, ny, nz, c=c)
        r = c / 0.5
        if r in (2.0, 0.5):
            assert np.array_equal(dc, r * d5)
        else:
            np.testing.assert_allclose(dc, r * d5, rtol=1e-14, atol=0)
        T5, Tc = O.elemental_types(V, d5, ph), O.elemental_types(V, dc, ph)
        np.testing.assert_array_equal(Tc[:, 0], T5[:, 0])
        scale = np.abs(T5).max()
        assert np.abs(Tc[:, 1] - r * r * T5[:, 1]).max() <= 1e-13 * scale
        assert np.abs(Tc[:, 2:] - r * T5[:, 2:]).max() <= 1e-13 * scale
        # the laws have teeth: type 1 does not scale as c, types 2-4 not as c^2
        assert np.abs(Tc[:, 1] - r * T5[:, 1]).max() > 1e-3 * scale
        assert np.abs(Tc[:, 2:] - r * r * T5[:, 2:]).max() > 1e-3 * scale
