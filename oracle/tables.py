"""Spin tables of the contraction engine, regenerated from closed-form gamma matrices.

TEST INFRASTRUCTURE ONLY (oracle): imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product path.

The reference ships its tables as generated C arrays (include/constants_{rho,rhopi,pipi,
sigma,sigmapi}.h) written by devTools/comb_*.cpp on top of an external gamma-matrix library.
This module restates the generator formulas:

* rho     devTools/comb_rho.cpp:19-49 (rhoj), :52-74 (compute_lp_rp), :96-180 (row loop)
* rhopi   devTools/comb_rhopi.cpp (rhoj, compute_ip, row loop over idir,iop,s1..s6)
* pipi    devTools/comb_pipi.cpp:51 (square), :92 (doubleTriangle), :105-123 (doubleTriangleHor)
* sigma   devTools/comb_sigma.cpp:50 (l,k loop, coef gamma5(s2,s3)*gamma5(s4,s1))
* sigmapi devTools/comb_sigmapi.cpp:70 (triangle), hor_fish (all 16 pairs, coef 1)

in the chiral gamma basis below, in which every generated array agrees bit for bit with the
reference headers (checked by SHA-256 digests in tests/test_tables.py).
"""
import hashlib

import numpy as np

_I = 1j
GAMMA = {
    1: np.array([[0, 0, 0, -_I], [0, 0, -_I, 0], [0, _I, 0, 0], [_I, 0, 0, 0]], dtype=complex),
    2: np.array([[0, 0, 0, 1], [0, 0, -1, 0], [0, -1, 0, 0], [1, 0, 0, 0]], dtype=complex),
    3: np.array([[0, 0, -_I, 0], [0, 0, 0, _I], [_I, 0, 0, 0], [0, -_I, 0, 0]], dtype=complex),
    4: np.array([[0, 0, -1, 0], [0, 0, 0, -1], [-1, 0, 0, 0], [0, -1, 0, 0]], dtype=complex),
    5: np.diag([1, 1, -1, -1]).astype(complex),
}
G5 = GAMMA[5]


def rhoj(j, i, r, c, dag=False):
    """Rho interpolator spin matrix element (devTools/comb_rho.cpp:19-49)."""
    g = GAMMA[i + 1]
    if j == 1 or j == 3:
        return -g[r, c] if dag else g[r, c]
    if j == 2:
        return (GAMMA[4] @ g)[r, c]
    if j == 4:
        res = GAMMA[4] @ GAMMA[4]
        return -0.5 * res[r, c] if dag else 0.5 * res[r, c]
    raise ValueError(j)


def compute_lp_rp(l, k, i):
    """Elemental types used at source (lp, from k) and sink (rp, from l); comb_rho.cpp:52-74."""
    sel = [0, 0, 1, 2 + i]
    return sel[k], sel[l]


def compute_ip(idir, iop):
    return [0, 0, 1, 2 + idir][iop]


def _c(z):
    z = complex(z)
    return z


def gen_rho():
    idx, lprp, coef = [], [], []
    for l in range(4):
        for k in range(4):
            for i in range(3):
                for s1 in range(4):
                    for s2 in range(4):
                        for s3 in range(4):
                            for s4 in range(4):
                                c = (rhoj(k + 1, i, s4, s2, True) * G5[s2, s2] * G5[s1, s1]
                                     * rhoj(l + 1, i, s1, s3))
                                if l * 4 + k < 12:
                                    c = -c
                                if abs(c) != 0:
                                    idx.append((s1, s2, s3, s4))
                                    lprp.append(compute_lp_rp(l, k, i))
                                    coef.append((c.real, c.imag))
    return dict(spinIndices_rho=idx, lp_rp_rho=lprp, coef_rho=coef)


def gen_rhopi():
    idx, ip, coef = [], [], []
    for idir in range(3):
        for iop in range(4):
            for s1 in range(4):
                for s2 in range(4):
                    for s3 in range(4):
                        for s4 in range(4):
                            for s5 in range(4):
                                for s6 in range(4):
                                    c = rhoj(iop + 1, idir, s4, s5) * G5[s3, s4] * G5[s6, s1]
                                    if abs(c) != 0:
                                        idx.append((s1, s2, s3, s4, s5, s6))
                                        ip.append(compute_ip(idir, iop))
                                        coef.append((c.real, c.imag))
    return dict(spinIndices_rhopi=idx, ip_rhopi=ip, coef_rhopi=coef)


def _six(cf):
    idx, coef = [], []
    for s1 in range(4):
        for s2 in range(4):
            for s3 in range(4):
                for s4 in range(4):
                    for s5 in range(4):
                        for s6 in range(4):
                            c = cf(s1, s2, s3, s4, s5, s6)
                            if abs(c) != 0:
                                idx.append((s1, s2, s3, s4, s5, s6))
                                coef.append((c.real, c.imag))
    return idx, coef


def gen_pipi():
    sq_i, sq_c = _six(lambda s1, s2, s3, s4, s5, s6: G5[s4, s5] * G5[s6, s1])
    dt_i, dt_c = _six(lambda s1, s2, s3, s4, s5, s6: G5[s2, s3] * G5[s6, s1])
    hor_i = [(s1, s2, s3, s4) for s1 in range(4) for s2 in range(4) for s3 in range(4)
             for s4 in range(4)]
    hor_c = [(1.0, 0.0)] * 256
    return dict(spinIndices_pipi_square=sq_i, coef_pipi_square=sq_c,
                spinIndices_pipi_doubleTriangle=dt_i, coef_pipi_doubleTriangle=dt_c,
                spinIndices_pipi_doubleTriangle_hor=hor_i, coef_pipi_doubleTriangle_hor=hor_c)


def gen_sigma():
    idx, lprp, coef = [], [], []
    for l in range(2):
        for k in range(2):
            for s1 in range(4):
                for s2 in range(4):
                    for s3 in range(4):
                        for s4 in range(4):
                            c = G5[s2, s3] * G5[s4, s1]
                            if abs(c) != 0:
                                idx.append((s1, s2, s3, s4))
                                lprp.append((l, k))
                                coef.append((c.real, c.imag))
    return dict(spinIndices_sigma=idx, lp_rp_sigma=lprp, coef_sigma=coef)


def gen_sigmapi():
    hf_i = [(s1, s2) for s1 in range(4) for s2 in range(4)]
    hf_c = [(1.0, 0.0)] * 16
    idx, ip, coef = [], [], []
    for ik in range(2):
        for s1 in range(4):
            for s2 in range(4):
                for s3 in range(4):
                    for s4 in range(4):
                        for s5 in range(4):
                            c = G5[s1, s2] * G5[s3, s4]
                            if abs(c) != 0:
                                idx.append((s1, s2, s3, s4, s5))
                                ip.append(ik)
                                coef.append((c.real, c.imag))
    return dict(spinIndices_sigmapi_hor_fish=hf_i, coef_sigmapi_hor_fish=hf_c,
                spinIndices_sigmapi_triangle=idx, ip_sigmapi_triangle=ip,
                coef_sigmapi_triangle=coef)


def all_tables():
    """Every table as numpy arrays: indices int16, coefficients float32 (as the reference)."""
    out = {}
    for gen in (gen_rho, gen_rhopi, gen_pipi, gen_sigma, gen_sigmapi):
        for name, rows in gen().items():
            if name.startswith("coef"):
                out[name] = np.array(rows, dtype=np.float32)
            else:
                out[name] = np.array(rows, dtype=np.int16)
    return out


def digest(arr):
    """SHA-256 of the little-endian row-major bytes; -0.0 coefficients normalised to +0.0."""
    a = np.ascontiguousarray(arr)
    if a.dtype == np.float32:
        a = a + np.float32(0.0)  # -0.0 + 0.0 == +0.0
        a = a.astype("<f4")
    else:
        a = a.astype("<i2")
    return hashlib.sha256(a.tobytes()).hexdigest()


# SHA-256 digests of the 20 reference arrays (include/constants_*.h), serialised as digest().
REFERENCE_DIGESTS = {
    "spinIndices_rho": ((768, 4), "ab01522ff2aa69cc71f863586d22595cbd2e327ed3e5c7f356c716b202061e42"),
    "lp_rp_rho": ((768, 2), "54378a89dfa2474dcf3307ca682cdca848730246230479d19fe2317a31615249"),
    "coef_rho": ((768, 2), "79b21a643b54320a575ddbf9616ee082c05b8e637aa48e19a2d4a8ba796a57b7"),
    "spinIndices_rhopi": ((768, 6), "d0438de37da452a750acfdc0eac008577295e77376306f503e1a887a35a355b9"),
    "ip_rhopi": ((768,), "f46685f4eb7fcdde77bd7df0c798b54a26cc7a2a59a304ba5e7a9dcdff89cf3d"),
    "coef_rhopi": ((768, 2), "fb5ee68f176e73a85ff0ed25fc050be202cb0da2706c50d7233c0c4bcba14ce0"),
    "spinIndices_pipi_square": ((256, 6), "33ed3e9398daf2c25afdc8b3d0f980460dcae480021af78aa168f23b86a993d6"),
    "coef_pipi_square": ((256, 2), "c4e074ca6b5c68c056b786b9a8a07db0365a7aec2574727c5250bff800a9a170"),
    "spinIndices_pipi_doubleTriangle": ((256, 6), "859f21c7e690de99fa6c0cadb56c21e205f48ade596bbee700c877fe6af5142d"),
    "coef_pipi_doubleTriangle": ((256, 2), "080f6384ddb1bb8492bab69e13313833288917f186c64de4d0b28b428a1824b4"),
    "spinIndices_pipi_doubleTriangle_hor": ((256, 4), "305ce64f42836e9bf230e74f1bb21f550e8d3ceb623f010557a2eee0287f2ae6"),
    "coef_pipi_doubleTriangle_hor": ((256, 2), "0da648fdcf61b5ea6383960583bdc74287c3fa758c6e5a7eb7d234f2d7c3283f"),
    "spinIndices_sigma": ((64, 4), "9037765098d11f02f09dbefdb36e6a8bfea126d168ed73b91069bcac4a6604cc"),
    "lp_rp_sigma": ((64, 2), "4cb275654c3ca21297ef088d936db0661c35ae51bb637765a46b01c7007af892"),
    "coef_sigma": ((64, 2), "36f6c8a3f2a39e56f3af95deb482ee19bc7d51cafea65a4594ba6c8db4e96223"),
    "spinIndices_sigmapi_hor_fish": ((16, 2), "a84070effa39477e479af670e1df92ca0718905a0490dc82d0d85119d289257f"),
    "coef_sigmapi_hor_fish": ((16, 2), "885e62971209622046096c40c50b2ccbd4cb790a6972d65d95034ebff38dcc89"),
    "spinIndices_sigmapi_triangle": ((128, 5), "967d4258aa82bcd739313b4e78fcdab122f59b699bcf20b809091dfd99dcb56d"),
    "ip_sigmapi_triangle": ((128,), "b3ac87cacb36f87cc2ee9d83b7c1dae0d162ec9229a88662d5eb9177a8a724f5"),
    "coef_sigmapi_triangle": ((128, 2), "ffdd927af5274daacdefe6fd69d2d7fe97bf82826a43ebad2bfb118da7e592c1"),
}
