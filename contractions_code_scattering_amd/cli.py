"""Command-line driver for the contraction flows, with the option names of the reference's drivers
(lib/catchOpts.cpp:45-66, tests/run_compute_*.sh).

The perambulator is read from --propname (the readPropH file format). The elementals either come
from --elemname or are built on the GPU from eigenvectors (and, except for pi-pi, gauge links):

* --elemname: complex double driver arrays, one record per momentum combination, in the order of
  the calculate* arguments:

  | diagram | record |
  |---|---|
  | rho | mom [5][Nt][Nvec][Nvec] |
  | rhopi | P1 [Nt][Nvec][Nvec], P2, P1P2 [5][Nt][Nvec][Nvec] |
  | pipi | P1, P2, P3, P4 |
  | sigma | mom [4][Nt][Nvec][Nvec] |
  | sigmapi | P1, P2, P1P2 [2][Nt][Nvec][Nvec] |

* from eigenvectors: --eigename (complex double [Nt][Nvec][Nz][Ny][Nx][3]) and --momname (the
  reference's momentum list: Nmom * 3 integers per line, momComb lines; with Nmom 2 the third
  momentum is p1 + p2, lib/catchOpts.cpp:197-209). The derivative types need --linkname: the
  smeared gauge links with their boundary conditions applied, complex double
  [Nt][3][Nz][Ny][Nx][3][3] (U_j(x), j = x, y, z). Each record follows its driver, with
  e(+-p) = exp(+-2 pi i p.x / L) and T_k(e) the five types of lc_build_elementals_types:

  | diagram | Nmom | record | driver |
  |---|---|---|---|
  | rho | 1 | T_0..4(e(-p1)) | tests/compute_rho.cpp:116-229 |
  | rhopi | 2 | T_0(e(-p1)), T_0(e(-p2)), T_0..4(e(+p3)) | tests/compute_rhopi.cpp:125-245 |
  | pipi | 4 | T_0(e(-p1)), T_0(e(-p2)), T_0(e(+p3)), T_0(e(+p4)) | tests/compute_pipi_I0.cpp:133-146 |
  | sigma | 1 | T_0(e(+p1)), T_0(e(-p1)), T_1(e(+p1)), T_1(e(-p1)) | tests/compute_sigma.cpp:113-195 |
  | sigmapi | 2 | T_0(e(-p1)), T_0(e(-p2)), T_0..1(e(+p3)) | tests/compute_sigmapi.cpp:115-180 |

  Elementals and covariant derivatives are formed on the GPU (lc_build_elementals,
  lc_covariant_derivative, lc_build_elementals_types).

Each momentum combination appends its text to --outname, as one call of the reference's
calculate* does. With --binary, a raw FP64 side channel is also written to <outname>.bin.

    python -m contractions_code_scattering_amd.cli pipi_I0 --propname prop.bin --eigename evec.bin \\
        --momname mom_pipi.dat --Nmom 4 --momComb 1 --Nx 16 --Ny 16 --Nz 16 --Nt 32 --Nvec 64 \\
        --outname pipi.dat
"""
import argparse
import ctypes
import os
import sys

import numpy as np

from . import _capi
from ._capi import check

DIAGRAMS = ("rho", "rhopi", "pipi_I0", "pipi_I1", "sigma", "sigmapi")
# elemental arrays of one record, in units of [Nt][Nvec][Nvec]
NMOM = {"rho": 1, "rhopi": 2, "pipi_I0": 4, "pipi_I1": 4, "sigma": 1, "sigmapi": 2}
RECORD = {"rho": [5], "rhopi": [1, 1, 5], "pipi_I0": [1, 1, 1, 1], "pipi_I1": [1, 1, 1, 1], "sigma": [4],
          "sigmapi": [1, 1, 2]}


def parse(argv=None):
    ap = argparse.ArgumentParser(prog="python -m contractions_code_scattering_amd.cli",
                                 description="LapH meson contractions on the GPU")
    ap.add_argument("diagram", choices=DIAGRAMS)
    ap.add_argument("--prec", choices=["double", "single"], default="double")
    ap.add_argument("--devID", type=int, default=0)
    ap.add_argument("--Nx", type=int, default=16)
    ap.add_argument("--Ny", type=int, default=16)
    ap.add_argument("--Nz", type=int, default=16)
    ap.add_argument("--Nt", type=int, default=32)
    ap.add_argument("--Nvec", type=int, default=100)
    ap.add_argument("--Nmom", type=int, default=1)
    ap.add_argument("--momComb", type=int, default=1)
    ap.add_argument("--idir", type=int, default=0)
    ap.add_argument("--propname", required=True)
    ap.add_argument("--eigename", default="")
    ap.add_argument("--momname", default="")
    ap.add_argument("--elemname", default="")
    ap.add_argument("--linkname", default="", help="smeared gauge links (derivative elemental types)")
    ap.add_argument("--covderivNorm", type=float, default=0.5,
                    help="normalisation c of the covariant derivative c [U v(x+j) - U^dag v(x-j)] (the drivers' "
                         "external covderiv; not in the reference): type 1 scales as c^2, types 2-4 as c")
    ap.add_argument("--outname", default="out.dat")
    ap.add_argument("--binary", action="store_true", help="also write <outname>.bin (FP64 values)")
    a = ap.parse_args(argv)
    if not a.elemname:
        if not (a.eigename and a.momname):
            ap.error("give --elemname, or --eigename and --momname")
        if not a.diagram.startswith("pipi") and not a.linkname:
            ap.error("%s from eigenvectors needs --linkname (its derivative elemental types)" % a.diagram)
        if a.Nmom != NMOM[a.diagram]:
            ap.error("%s needs --Nmom %d" % (a.diagram, NMOM[a.diagram]))
    if a.idir not in (0, 1, 2):
        ap.error("--idir must be 0, 1 or 2")
    return a


def read_momenta(path, nmom, ncomb):
    """The reference's momentum list (lib/catchOpts.cpp:106-200): ncomb lines of nmom * 3 integers."""
    rows = [ln.split() for ln in open(path).read().splitlines() if ln.strip()]
    if len(rows) != ncomb:
        raise SystemExit("Number of lines in the file are not in agreement with the number of momentum "
                         "combinations you passed")
    out = []
    for r in rows:
        if len(r) != 3 * nmom:
            raise SystemExit("Number of columns in the file are not in agreement with the number of momenta "
                             "you passed")
        v = [int(x) for x in r]
        out.append([tuple(v[3 * i:3 * i + 3]) for i in range(nmom)])
    return out


def site_phases(moms, signs, nx, ny, nz):
    """exp(sign * 2 pi i p.x / L) for each (p, sign), over sites x + Nx (y + Ny z) and 3 colours."""
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    out = []
    for (px, py, pz), s in zip(moms, signs):
        arg = 2 * np.pi * (x * px / nx + y * py / ny + z * pz / nz)
        out.append(np.repeat(np.exp(1j * s * arg).ravel(), 3))
    return np.ascontiguousarray(np.stack(out))


def pipi_elementals(evec_dev, moms, a):
    from .engine import build_elementals
    from .devmem import to_device
    ph = to_device(site_phases(moms, [-1, -1, +1, +1], a.Nx, a.Ny, a.Nz))
    phi = build_elementals(evec_dev, ph).to_numpy()  # [4][Nt][Nvec][Nvec]
    return [np.ascontiguousarray(phi[i]) for i in range(4)]


def driver_elementals(evec_dev, dvec_dev, moms, a):
    """The record of one momentum combination for rho, rhopi, sigma or sigmapi, as its driver forms
    it (table in the module docstring); moms are the combination's Nmom momenta."""
    from .engine import build_elementals_types
    from .devmem import to_device

    def types(plist, signs):  # [len(plist)][5][Nt][Nvec][Nvec]
        ph = to_device(site_phases(plist, signs, a.Nx, a.Ny, a.Nz))
        return build_elementals_types(evec_dev, dvec_dev, ph).to_numpy()
    c = np.ascontiguousarray
    d = a.diagram
    if d == "rho":
        return [c(types([moms[0]], [-1])[0])]
    if d == "sigma":
        T = types([moms[0], moms[0]], [+1, -1])
        return [c(np.stack([T[0, 0], T[1, 0], T[0, 1], T[1, 1]]))]
    p3 = tuple(x + y for x, y in zip(moms[0], moms[1]))
    T = types([moms[0], moms[1], p3], [-1, -1, +1])
    last = T[2] if d == "rhopi" else T[2, :2]
    return [c(T[0, 0]), c(T[1, 0]), c(last)]


def _p(x):
    return ctypes.c_void_p(x.ctypes.data)


def main(argv=None):
    a = parse(argv)
    L = _capi.lib()
    nt, nv = a.Nt, a.Nvec
    prec = 1 if a.prec == "double" else 0
    check(L.lc_contract_init(nv, nt, a.devID))
    try:
        check(L.lc_set_binary_output(1 if a.binary else 0))
        check(L.lc_set_i_dir(a.idir))
        prop = ctypes.c_void_p()
        check(L.lc_read_prop(a.propname.encode(), prec, ctypes.byref(prop)))
        try:
            blk = nt * nv * nv
            records = []
            if a.elemname:
                n_per = sum(RECORD[a.diagram]) * blk
                raw = np.fromfile(a.elemname, dtype="<c16")
                if raw.size != n_per * a.momComb:
                    raise SystemExit("%s holds %d complex values, expected %d (%d momentum combinations)"
                                     % (a.elemname, raw.size, n_per * a.momComb, a.momComb))
                for r in range(a.momComb):
                    rec = raw[r * n_per:(r + 1) * n_per]
                    arrs, o = [], 0
                    for k in RECORD[a.diagram]:
                        arrs.append(np.ascontiguousarray(rec[o:o + k * blk]))
                        o += k * blk
                    records.append(arrs)
            else:
                from .devmem import to_device
                moms = read_momenta(a.momname, a.Nmom, a.momComb)
                nsite = 3 * a.Nx * a.Ny * a.Nz
                ev = np.fromfile(a.eigename, dtype="<c16")
                if ev.size != nt * nv * nsite:
                    raise SystemExit("%s holds %d complex values, expected Nt*Nvec*3*Nx*Ny*Nz = %d"
                                     % (a.eigename, ev.size, nt * nv * nsite))
                evec = to_device(ev.reshape(nt, nv, nsite))
                if a.diagram.startswith("pipi"):
                    for m in moms:
                        records.append(pipi_elementals(evec, m, a))
                else:
                    from .engine import covariant_derivative
                    nlink = nt * 3 * (nsite // 3) * 9
                    lk = np.fromfile(a.linkname, dtype="<c16")
                    if lk.size != nlink:
                        raise SystemExit("%s holds %d complex values, expected Nt*3*Nx*Ny*Nz*9 = %d"
                                         % (a.linkname, lk.size, nlink))
                    links = to_device(lk.reshape(nt, 3, nsite // 3, 3, 3))
                    dvec = covariant_derivative(links, evec, a.Nx, a.Ny, a.Nz, c=a.covderivNorm)
                    del links
                    for m in moms:
                        records.append(driver_elementals(evec, dvec, m, a))
                    del dvec
                del evec
            for arrs in records:
                ptrs = [_p(x) for x in arrs]
                out = a.outname.encode()
                if a.diagram == "rho":
                    check(L.lc_calculate_rho(nv, nt, prop, *ptrs, prec, out))
                elif a.diagram == "rhopi":
                    check(L.lc_calculate_rhopi(a.idir, nv, nt, prop, *ptrs, prec, out))
                elif a.diagram.startswith("pipi"):
                    check(L.lc_calculate_pipi(nv, nt, prop, *ptrs, prec, out, 0 if a.diagram == "pipi_I0" else 1))
                elif a.diagram == "sigma":
                    check(L.lc_calculate_sigma(nv, nt, prop, *ptrs, prec, out))
                else:
                    check(L.lc_calculate_sigmapi(nv, nt, prop, *ptrs, prec, out))
        finally:
            L.lc_free_host(prop)
    finally:
        check(L.lc_contract_destroy())
    return 0


if __name__ == "__main__":
    sys.exit(main())
