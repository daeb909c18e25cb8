"""Seeded synthetic inputs in the reference's layouts.

* perambulator tau(t, t0): host layout [Nt][Nt][4][4][Nv][Nv] (lib/contract.cpp:140,152); as a
  D x D matrix (D = 4 Nv) its entries are i.i.d. CN(0,1) * exp(-m |t - t0|) / sqrt(D).
* elementals Phi(p, t): [Nt][Nv][Nv]; Hermitian (G + G^dag) / (2 sqrt(Nv)) for zero momentum,
  or i.i.d. CN(0,1)/sqrt(Nv) with Phi(-p, t) = Phi(p, t)^dag.

CN(0,1) = (N(0,1/2) + i N(0,1/2)). Generators are numpy PCG64 streams, so identical seeds give
identical inputs on every machine.
"""
import numpy as np


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def cn(g, shape, dtype=np.complex128):
    if np.dtype(dtype) == np.complex128:
        # the same values as below, built in place (a few GB smaller peak for large perambulators)
        out = np.empty(shape, np.complex128)
        out.real = g.standard_normal(shape)
        out.imag = g.standard_normal(shape)
        out *= np.sqrt(0.5)
        return out
    re = g.standard_normal(shape)
    im = g.standard_normal(shape)
    return ((re + 1j * im) * np.sqrt(0.5)).astype(dtype)


def perambulator(g, nt, nv, mass, dtype=np.complex128):
    """tau[t, t0, s1, s2, v1, v2]."""
    d = 4 * nv
    tau = cn(g, (nt, nt, 4, 4, nv, nv), dtype)
    t = np.arange(nt)
    decay = np.exp(-mass * np.abs(t[:, None] - t[None, :])) / np.sqrt(d)
    tau *= decay[:, :, None, None, None, None].astype(tau.real.dtype)
    return tau


def hermitian_elemental(g, nt, nv, dtype=np.complex128):
    G = cn(g, (nt, nv, nv), dtype)
    return ((G + np.conj(np.swapaxes(G, -1, -2))) / (2 * np.sqrt(nv))).astype(dtype)


def random_elemental(g, nt, nv, dtype=np.complex128):
    return (cn(g, (nt, nv, nv), dtype) / np.sqrt(nv)).astype(dtype)


def dagger(m):
    return np.ascontiguousarray(np.conj(np.swapaxes(m, -1, -2)))


def diag_of(tau):
    nt = tau.shape[0]
    return np.ascontiguousarray(tau[np.arange(nt), np.arange(nt)])
