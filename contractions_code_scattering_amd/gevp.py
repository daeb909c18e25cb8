"""Correlator matrices across operators and frames, and their generalized eigenvalue problem.

The reference's drivers stop at per-diagram text: dumpData averages each diagram over the
sources and prints it with its factor (lib/contract.cpp:616-1053). The next analysis step joins
diagrams into a matrix of correlators C_ij(t) = <O_i(t) O_j(0)^dag> between several operators of
one frame (total momentum, irrep) and solves the generalized eigenvalue problem (GEVP) per time.
This module does that on the host from the engines' outputs (numpy / scipy):

* ``source_average``: C(dt) = mean over the sources t0 of c(tf = (t0 + dt) mod Nt, t0), dumpData's
  average (lib/contract.cpp:620-632) restricted to a source list.
* The reference's diagram factors (``PIPI_WEIGHTS``, ``RHOPI_WEIGHT``, ``SIGMA_WEIGHT``,
  ``SIGMAPI_TRIANGLE_WEIGHT``), and ``pipi`` / ``rho_pipi_diagrams`` / ``sigma_set_diagrams`` that
  apply them to the engines' unweighted outputs.
* ``Frame`` / ``assemble_frames``: each matrix element is a weighted sum of named diagram arrays;
  an element given for (i, j) only is completed by hermiticity, C_ji(t) = conj(C_ij(t)).
* ``gevp``: C(t) v_n = lambda_n(t, t0) C(t0) v_n for every t, with C(t0) Hermitian positive
  definite; eigenvalues ordered so that n = 0 is the lowest energy.
* ``effective_energies``: E_n(t) = log(lambda_n(t) / lambda_n(t + dt)) / dt.
"""
import math
from dataclasses import dataclass, field

import numpy as np

# pi-pi isospin combinations of the unweighted diagrams, as run_ContractPiPi forms them:
# I = 1 lib/contract_kernels_pipi.cu:288-303, I = 0 :463-478
PIPI_WEIGHTS = {
    1: {"square": -2.0, "doubleTriangle": 2.0, "star": -1.0, "fish": 1.0},
    0: {"square": -1.0 / 3.0, "doubleTriangle": -1.0 / 3.0, "star": 1.0, "fish": 1.0,
        "doubleTriangleHor": -5.0 / 3.0},
}
RHOPI_WEIGHT = 2.0                               # rho-pi dumpData (lib/contract.cpp:722)
SIGMA_WEIGHT = -1.0                              # sigma "twop" dumpData (lib/contract.cpp:942)
SIGMAPI_TRIANGLE_WEIGHT = -2.0 / math.sqrt(6.0)  # sigma-pi triangle dumpData (lib/contract.cpp:1045)

# diagram order of the rho + pi-pi engine's corr / avg arrays
RHO_PIPI_NAMES = ("rho_x", "rho_y", "rho_z", "square", "doubleTriangle", "doubleTriangleHor", "star", "fish")
# per-source slots of the sigma-set engine's raw array
SIGMA_SET_NAMES = ("sigma_00", "sigma_01", "sigma_10", "sigma_11", "triangle_0", "triangle_1",
                   "rhopi_0", "rhopi_1", "rhopi_2", "rhopi_3")


def source_average(per_source, src, nt):
    """per_source [..., nsrc, nt] indexed by (source k, sink tf), sources t0 = src[k]; returns
    [..., nt] indexed by dt: the mean over k of per_source[..., k, (src[k] + dt) % nt]."""
    c = np.asarray(per_source)
    src = list(src)
    if c.shape[-2] != len(src) or c.shape[-1] != nt:
        raise ValueError("per_source must be [..., %d sources, %d times]" % (len(src), nt))
    dt = np.arange(nt)
    out = np.zeros(c.shape[:-2] + (nt,), dtype=np.result_type(c.dtype, np.complex128))
    for k, t0 in enumerate(src):
        out += c[..., k, (t0 + dt) % nt]
    return out / len(src)


def rho_pipi_diagrams(avg):
    """The rho + pi-pi engine's source average avg [..., 8, nt] as a dict name -> [..., nt]."""
    a = np.asarray(avg)
    if a.shape[-2] != len(RHO_PIPI_NAMES):
        raise ValueError("avg must be [..., 8, nt]")
    return {n: a[..., i, :] for i, n in enumerate(RHO_PIPI_NAMES)}


def pipi(avg, iso):
    """Isospin-`iso` pi-pi correlator [..., nt] from the engine's unweighted avg [..., 8, nt]."""
    if iso not in PIPI_WEIGHTS:
        raise ValueError("iso must be 0 or 1")
    d = rho_pipi_diagrams(avg)
    return sum(w * d[name] for name, w in PIPI_WEIGHTS[iso].items())


def sigma_set_diagrams(raw, src, nt, weighted=True):
    """The sigma-set engine's per-source raw [..., nsrc, nt, 10], source-averaged per slot into a
    dict name -> [..., nt]; weighted=True applies the reference's factors (sigma -1, sigma-pi
    triangle -2/sqrt(6), rho-pi 2)."""
    r = np.asarray(raw)
    if r.shape[-1] != len(SIGMA_SET_NAMES):
        raise ValueError("raw must be [..., nsrc, nt, 10]")
    out = {}
    for i, n in enumerate(SIGMA_SET_NAMES):
        c = source_average(r[..., i], src, nt)
        if weighted:
            c = c * (SIGMA_WEIGHT if n.startswith("sigma") else
                     SIGMAPI_TRIANGLE_WEIGHT if n.startswith("triangle") else RHOPI_WEIGHT)
        out[n] = c
    return out


@dataclass
class Frame:
    """Operators of one frame and how each matrix element is formed: elements[(i, j)] is a list of
    (weight, key) terms, key naming an array [nt] in the data handed to assemble_frames."""
    name: str
    operators: list
    elements: dict = field(default_factory=dict)

    def add(self, i, j, *terms):
        self.elements.setdefault((i, j), []).extend(terms)
        return self


def assemble(elements, n):
    """elements {(i, j): array [nt]} -> C [nt][n][n] (complex). An element given only as (i, j)
    supplies C_ji = conj(C_ij); every pair needs at least one of the two."""
    nt = None
    for v in elements.values():
        nt = np.asarray(v).shape[-1]
        break
    if nt is None:
        raise ValueError("no elements")
    C = np.zeros((nt, n, n), np.complex128)
    for i in range(n):
        for j in range(n):
            if (i, j) in elements:
                v = np.asarray(elements[(i, j)], np.complex128)
            elif (j, i) in elements:
                v = np.conj(np.asarray(elements[(j, i)], np.complex128))
            else:
                raise ValueError("matrix element (%d, %d) is missing (and so is (%d, %d))" % (i, j, j, i))
            if v.shape != (nt,):
                raise ValueError("element (%d, %d) has shape %s, expected (%d,)" % (i, j, v.shape, nt))
            C[:, i, j] = v
    return C


def assemble_frames(frames, data):
    """{frame name: C [nt][n][n]} with C_ij = sum of weight * data[key] over the element's terms."""
    out = {}
    for f in frames:
        elems = {}
        for ij, terms in f.elements.items():
            if not terms:
                raise ValueError("frame %s element %s has no terms" % (f.name, ij))
            elems[ij] = sum(w * np.asarray(data[key]) for w, key in terms)
        out[f.name] = assemble(elems, len(f.operators))
    return out


def gevp(C, t0):
    """Generalized eigenvalues and vectors of C [nt][n][n] against C(t0): C(t) v = lambda C(t0) v.

    C is made exactly Hermitian first ((C + C^dag) / 2, which removes only statistical or rounding
    asymmetry); C(t0) must be positive definite. Returns lam [nt][n] and the eigenvectors
    v [nt][n][n] (columns, normalised to v^dag C(t0) v = 1). For t > t0 the eigenvalues are in
    descending order and for t < t0 in ascending order, so that in the noiseless case index n is
    the n-th lowest energy at every t (lambda_n = exp(-E_n (t - t0)) falls with t for t > t0 and
    exceeds 1 for t < t0)."""
    from scipy.linalg import eigh
    C = np.asarray(C, np.complex128)
    if C.ndim != 3 or C.shape[1] != C.shape[2]:
        raise ValueError("C must be [nt][n][n]")
    nt, n = C.shape[0], C.shape[1]
    if not 0 <= t0 < nt:
        raise ValueError("t0 out of range")
    H = 0.5 * (C + np.conj(np.swapaxes(C, 1, 2)))
    B = H[t0]
    if np.linalg.eigvalsh(B).min() <= 0:
        raise ValueError("C(t0) is not positive definite")
    lam = np.zeros((nt, n))
    vec = np.zeros((nt, n, n), np.complex128)
    for t in range(nt):
        w, v = eigh(H[t], B)
        order = np.argsort(w) if t < t0 else np.argsort(w)[::-1]
        lam[t] = w[order]
        vec[t] = v[:, order]
    return lam, vec


def effective_energies(lam, dt=1):
    """E_n(t) = log(lambda_n(t) / lambda_n(t + dt)) / dt, [nt - dt][n]; NaN where the ratio is not
    positive."""
    lam = np.asarray(lam, np.float64)
    r = lam[:-dt] / lam[dt:]
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(r > 0, np.log(np.where(r > 0, r, 1.0)) / dt, np.nan)
