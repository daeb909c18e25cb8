"""Oracle for the LapH meson contraction engine.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker. Nothing in the product package imports it.

Four independent pieces:

* ``oracle/_build/liboracle.so`` (oracle.cpp): a row-by-row restatement of the reference's
  CUDA cores and host launchers, one table row at a time, FP64, OpenMP.
* The reference's host flows around those launchers (lib/interface.cpp, lib/contract.cpp),
  restated in numpy below: elemental transposes, the per-sink-row loops, the source average of
  ``dumpData`` and its ``%+e`` text output, and the perambulator file permutation of
  ``readPropH``.
* ``covariant_derivative`` / ``elemental_types``: the drivers' elemental construction
  (tests/compute_rho.cpp:116-229) from links and eigenvectors. The derivative stands in for
  gwu-qcd's covderiv, so its normalisation is parity unpinned against gwu-qcd; the rest is pinned
  by gauge covariance and free-field closed forms.
* ``dense``: a brute-force evaluator that builds explicit 4Nv x 4Nv matrices and evaluates
  every diagram as a trace of gamma-matrix / elemental / perambulator products in spin x
  eigenvector space. It shares no code with the row restatement, so agreement between the two
  pins the table semantics (operand order, conj placement, spin-block orientation).
"""
import ctypes
import os

import numpy as np

from . import tables as _tables

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None
TABLES = _tables.all_tables()


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _cc(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a


def num_threads():
    return lib().orc_num_threads()


# ------------------------------------------------------------------ per-row launchers (C++)
def rho_row(row, mom5, tf):
    """run_ContractRho restated: row [Nt,4,4,Nv,Nv], mom5 [Nt,5,Nv,Nv] -> [Nt,16,3]."""
    row, mom5 = _cc(row), _cc(mom5)
    nt, nv = row.shape[0], row.shape[-1]
    out = np.zeros((nt, 16, 3), np.complex128)
    T = TABLES
    lib().orc_rho_row(_p(row), _p(mom5), tf, nt, nv, _p(T["spinIndices_rho"]),
                      _p(T["lp_rp_rho"]), _p(T["coef_rho"]), _p(out))
    return out


def rho_row_range(row, mom5, tf, r0, r1):
    """rho_row restricted to spin-table rows [r0, r1) (rows 0..47 = op pair (l,k) = (0,0))."""
    row, mom5 = _cc(row), _cc(mom5)
    nt, nv = row.shape[0], row.shape[-1]
    out = np.zeros((nt, 16, 3), np.complex128)
    T = TABLES
    lib().orc_rho_row_range(_p(row), _p(mom5), tf, nt, nv, _p(T["spinIndices_rho"]),
                            _p(T["lp_rp_rho"]), _p(T["coef_rho"]), r0, r1, _p(out))
    return out


def rhopi_row(row, diag, p1, p2, p12x5, tf, idir):
    row, diag, p1, p2, p12x5 = map(_cc, (row, diag, p1, p2, p12x5))
    nt, nv = row.shape[0], row.shape[-1]
    out = np.zeros((nt, 4), np.complex128)
    T = TABLES
    lib().orc_rhopi_row(_p(row), _p(diag), _p(p1), _p(p2), _p(p12x5), tf, nt, nv, idir,
                        _p(T["spinIndices_rhopi"]), _p(T["ip_rhopi"]), _p(T["coef_rhopi"]),
                        _p(out))
    return out


def pipi_row(iso, row, diag, m1, m2, m3, m4, tf):
    """iso: 1 for I=1 (run_ContractPiPi), 0 for I=0 (run_ContractPiPi_I0). -> [Nt,5]."""
    row, diag, m1, m2, m3, m4 = map(_cc, (row, diag, m1, m2, m3, m4))
    nt, nv = row.shape[0], row.shape[-1]
    out = np.zeros((nt, 5), np.complex128)
    T = TABLES
    lib().orc_pipi_row(iso, _p(row), _p(diag), _p(m1), _p(m2), _p(m3), _p(m4), tf, nt, nv,
                       _p(T["spinIndices_pipi_square"]), _p(T["coef_pipi_square"]),
                       _p(T["spinIndices_pipi_doubleTriangle"]), _p(T["coef_pipi_doubleTriangle"]),
                       _p(T["spinIndices_pipi_doubleTriangle_hor"]),
                       _p(T["coef_pipi_doubleTriangle_hor"]), _p(out))
    return out


def loop_unity(diag, mom):
    diag, mom = _cc(diag), _cc(mom)
    nt, nv = diag.shape[0], diag.shape[-1]
    out = np.zeros((nt,), np.complex128)
    lib().orc_loop_unity(_p(diag), _p(mom), nt, nv, _p(out))
    return out


def sigma_row(row, x2, tf):
    row, x2 = _cc(row), _cc(x2)
    nt, nv = row.shape[0], row.shape[-1]
    out = np.zeros((nt, 4), np.complex128)
    T = TABLES
    lib().orc_sigma_row(_p(row), _p(x2), tf, nt, nv, _p(T["spinIndices_sigma"]),
                        _p(T["lp_rp_sigma"]), _p(T["coef_sigma"]), _p(out))
    return out


def sigmapi_fish_hor(diag, p1, p2):
    diag, p1, p2 = map(_cc, (diag, p1, p2))
    nt, nv = diag.shape[0], diag.shape[-1]
    out = np.zeros((nt,), np.complex128)
    T = TABLES
    lib().orc_sigmapi_fish_hor(_p(diag), _p(p1), _p(p2), nt, nv,
                               _p(T["spinIndices_sigmapi_hor_fish"]),
                               _p(T["coef_sigmapi_hor_fish"]), _p(out))
    return out


def sigmapi_triangle_row(row, diag, p1, p2, x2, tf):
    row, diag, p1, p2, x2 = map(_cc, (row, diag, p1, p2, x2))
    nt, nv = row.shape[0], row.shape[-1]
    out = np.zeros((nt, 2), np.complex128)
    T = TABLES
    lib().orc_sigmapi_triangle_row(_p(row), _p(diag), _p(p1), _p(p2), _p(x2), tf, nt, nv,
                                   _p(T["spinIndices_sigmapi_triangle"]),
                                   _p(T["ip_sigmapi_triangle"]), _p(T["coef_sigmapi_triangle"]),
                                   _p(out))
    return out


# ------------------------------------------------------------------ host flows (numpy)
def diag_of(tau):
    """copyToDeviceExtraDiag (lib/contract.cpp:549-570): tau(t,t) -> [Nt,4,4,Nv,Nv]."""
    nt = tau.shape[0]
    return np.ascontiguousarray(tau[np.arange(nt), np.arange(nt)])


def source_average(h):
    """dumpData average (e.g. lib/contract.cpp:619-631): C(dt) = 1/Nt sum_ti h[(ti+dt)%Nt][ti]."""
    nt = h.shape[0]
    out = np.zeros((nt,) + h.shape[2:], np.complex128)
    for dt in range(nt):
        for it in range(nt):
            out[dt] += h[(it + dt) % nt, it]
        out[dt] /= nt
    return out


def _e(z):
    return "%+e %+e" % (z.real, z.imag)


def text_rho(corr):
    """Contract_rho_rho::dumpData text (lib/contract.cpp:633-639)."""
    nt = corr.shape[0]
    lines = []
    for iop in range(16):
        for it in range(nt):
            c = corr[it, iop]
            lines.append("%d %d \t %s \t %s \t %s\n" % (iop, it, _e(c[0]), _e(c[1]), _e(c[2])))
    lines.append("\n")
    return "".join(lines)


def calculate_rho(tau, mom):
    """calculateRho (lib/interface.cpp:7-57) + contractRho (lib/contract.cpp:589-643).

    tau: host layout [Nt,Nt,4,4,Nv,Nv]; mom: driver layout [5,Nt,Nv,Nv].
    Returns (h_elem [Nt,Nt,16,3], averaged corr [Nt,16,3], text).
    """
    nt = tau.shape[0]
    mom5 = np.ascontiguousarray(np.transpose(mom, (1, 0, 2, 3)))
    h = np.stack([rho_row(tau[tf], mom5, tf) for tf in range(nt)])
    corr = source_average(h)
    return h, corr, text_rho(corr)


def calculate_rhopi(idir, tau, p1, p2, p12, i_dir_print=None):
    """calculateRhoPi (lib/interface.cpp:59-119) + contractRhoPi/dumpData (lib/contract.cpp:664-728).

    The text prints the process-global contract::i_dir, which the driver sets to idir."""
    nt = tau.shape[0]
    p12x5 = np.ascontiguousarray(np.transpose(p12, (1, 0, 2, 3)))
    diag = diag_of(tau)
    h = np.stack([rhopi_row(tau[tf], diag, p1, p2, p12x5, tf, idir) for tf in range(nt)])
    corr = source_average(h)
    pd = idir if i_dir_print is None else i_dir_print
    lines = []
    for iop in range(4):
        for it in range(nt):
            lines.append("%d %d %d \t %s\n" % (iop, pd, it, _e(2 * corr[it, iop])))
    lines.append("\n")
    return h, corr, "".join(lines)


def calculate_pipi(tau, p1, p2, p3, p4, iso):
    """calculatePiPi (lib/interface.cpp:122-209) + contractPiPi/dumpData (lib/contract.cpp:747-839).
    iso: 1 (I1) or 0 (I0)."""
    nt = tau.shape[0]
    diag = diag_of(tau)
    h = np.stack([pipi_row(iso, tau[tf], diag, p1, p2, p3, p4, tf) for tf in range(nt)])
    corr = source_average(h)
    names = ["Square", "DoubleTriangle", "Star", "Fish"] + (["DoubleTriangle_Hor"] if iso == 0 else [])
    lines = []
    for i, nm in enumerate(names):
        for it in range(nt):
            lines.append("%s: %d \t %s\n" % (nm, it, _e(corr[it, i])))
        lines.append("\n")
    return h, corr, "".join(lines)


def _loop_text(loop):
    return "".join("loop: %d \t %s\n" % (it, _e(z)) for it, z in enumerate(loop)) + "\n"


def calculate_sigma(tau, mom4):
    """calculateSigma (lib/interface.cpp:212-291): four loops over types 0..3, then sigma-sigma on
    types {1,3} (:253-259), output -corr (lib/contract.cpp:940-945)."""
    nt = tau.shape[0]
    diag = diag_of(tau)
    loops = [loop_unity(diag, mom4[i]) for i in range(4)]
    x2 = np.ascontiguousarray(np.stack([mom4[1], mom4[3]], axis=1))  # [t][2][v][v]
    h = np.stack([sigma_row(tau[tf], x2, tf) for tf in range(nt)])
    corr = source_average(h)
    text = "".join(_loop_text(l) for l in loops)
    lines = []
    for iop in range(4):
        for it in range(nt):
            lines.append("twop: %d %d \t %s\n" % (iop, it, _e(-corr[it, iop])))
    lines.append("\n")
    return loops, h, corr, text + "".join(lines)


def calculate_sigmapi(tau, p1, p2, p12):
    """calculateSigmaPi (lib/interface.cpp:294-390) + contractSigmaPi (lib/contract.cpp:970-1053).
    p12: driver layout [2,Nt,Nv,Nv]."""
    nt = tau.shape[0]
    diag = diag_of(tau)
    loops = [loop_unity(diag, p12[i]) for i in range(2)]
    x2 = np.ascontiguousarray(np.transpose(p12, (1, 0, 2, 3)))
    fh = sigmapi_fish_hor(diag, p1, p2)
    h = np.stack([sigmapi_triangle_row(tau[tf], diag, p1, p2, x2, tf) for tf in range(nt)])
    corr = source_average(h)
    text = "".join(_loop_text(l) for l in loops)
    text += "".join("fish_hor: %d \t %s\n" % (it, _e(z)) for it, z in enumerate(fh)) + "\n"
    f = -2.0 / np.sqrt(6.0)
    lines = []
    for iop in range(2):
        for it in range(nt):
            lines.append("triangle: %d %d \t %s\n" % (iop, it, _e(f * corr[it, iop])))
    lines.append("\n")
    return loops, fh, h, corr, text + "".join(lines)


def read_prop_file(path, nt, nv):
    """readPropH (lib/contract.cpp:91-160): header then complex128 [s1][v1][t1][s2][v2][t2];
    returns the host layout [t1][t2][s1][s2][v1][v2]."""
    with open(path, "rb") as f:
        f.read(100)
        hdr = np.frombuffer(f.read(8), dtype="<u4")
        if int(hdr[1]) != nv or int(hdr[0]) != nt:
            raise ValueError("header mismatch")
        body = np.frombuffer(f.read(), dtype="<c16")
    if body.size != nt * nt * 16 * nv * nv:
        raise ValueError("short read")
    b = body.reshape(4, nv, nt, 4, nv, nt)
    return np.ascontiguousarray(b.transpose(2, 5, 0, 3, 1, 4))


# ------------------------------------------------------------------ elemental types
def covariant_derivative(U, V, nx, ny, nz, c=0.5):
    """(nabla_j v)(x) = c [U_j(x) v(x + j) - U_j(x - j)^dag v(x - j)], periodic, j = x, y, z: the
    symmetric lattice covariant derivative standing in for gwu-qcd's lapmult::covderiv, which the
    drivers call (tests/compute_rho.cpp:140-141) and which is not part of the reference; its
    normalisation c (1/2: the central difference) is therefore not pinned to gwu-qcd.
    U [nt][3][nsite][3][3], V [nt][nv][3 nsite] with sites x + nx (y + ny z), colour fastest;
    returns [3][nt][nv][3 nsite]."""
    nt, nv = V.shape[:2]
    v = V.reshape(nt, nv, nz, ny, nx, 3)
    u = U.reshape(nt, 3, nz, ny, nx, 3, 3)
    out = np.empty((3, nt, nv, nz, ny, nx, 3), complex)
    for j, ax in enumerate((4, 3, 2)):          # x, y, z axes of v; of u[:, j] they are ax - 1
        uj = u[:, j]
        ub = np.roll(uj, 1, axis=ax - 1)        # U_j(x - j)
        fwd = np.einsum("tzyxab,tczyxb->tczyxa", uj, np.roll(v, -1, axis=ax))
        bwd = np.einsum("tzyxba,tczyxb->tczyxa", ub.conj(), np.roll(v, 1, axis=ax))
        out[j] = c * (fwd - bwd)
    return out.reshape(3, nt, nv, -1)


def elemental_types(V, dV, ph):
    """The drivers' five momentum-matrix types per momentum, tmpMom [type][t][r][c]
    (tests/compute_rho.cpp:116-229; cscalar_product(a, b) = sum conj(a) b):
    0 plain V^dag ph V (:116-129), 1 smeared sum_j dV_j^dag ph dV_j (:131-162),
    2 + j commutator V^dag ph dV_j - dV_j^dag ph V (:164-202). Returns [np][5][nt][nv][nv]."""
    def bil(X, Y):
        return np.einsum("trk,pk,tck->ptrc", X.conj(), ph, Y)
    out = np.empty((ph.shape[0], 5) + V.shape[:2] + (V.shape[1],), complex)
    out[:, 0] = bil(V, V)
    out[:, 1] = sum(bil(dV[j], dV[j]) for j in range(3))
    for j in range(3):
        out[:, 2 + j] = bil(V, dV[j]) - bil(dV[j], V)
    return out


# ------------------------------------------------------------------ dense brute force
class dense:
    """Explicit (4Nv)x(4Nv) spin x eigenvector matrices; index (s, v) -> s*Nv + v."""

    @staticmethod
    def tau(blocks):
        """[4,4,Nv,Nv] spin blocks -> dense D x D."""
        nv = blocks.shape[-1]
        return blocks.transpose(0, 2, 1, 3).reshape(4 * nv, 4 * nv)

    @staticmethod
    def kron(m):
        return np.kron(np.eye(4), m)

    @staticmethod
    def gam(g, nv):
        return np.kron(g, np.eye(nv))

    @staticmethod
    def w_hor(blocks):
        """W_{ab} = conj(P_{ba}) with no transpose of the eigenvector indices."""
        return dense.tau(np.conj(blocks.transpose(1, 0, 2, 3)))

    @staticmethod
    def rho(tau_b, mom5, tf, ti):
        """[16,3]: rho(l,k,i) = -+ Tr[B^dag (g5 L) A (K g5)], B = M_rp(tf) tau, A = tau M_lp(ti)."""
        nv = tau_b.shape[-1]
        T = dense.tau(tau_b)
        g5 = dense.gam(_tables.G5, nv)
        out = np.zeros((16, 3), complex)
        for l in range(4):
            for k in range(4):
                for i in range(3):
                    lp, rp = _tables.compute_lp_rp(l, k, i)
                    L = np.array([[_tables.rhoj(l + 1, i, r, c) for c in range(4)] for r in range(4)])
                    K = np.array([[_tables.rhoj(k + 1, i, r, c, True) for c in range(4)] for r in range(4)])
                    A = T @ dense.kron(mom5[ti, lp])
                    B = dense.kron(mom5[tf, rp]) @ T
                    v = np.trace(B.conj().T @ g5 @ dense.gam(L, nv) @ A @ dense.gam(K, nv) @ g5)
                    out[l * 4 + k, i] = -v if l * 4 + k < 12 else v
        return out

    @staticmethod
    def rhopi(tau_b, diag_b, p1, p2, p12x5, tf, ti, idir):
        nv = tau_b.shape[-1]
        T = dense.tau(tau_b)
        Dm = dense.tau(diag_b)
        g5 = dense.gam(_tables.G5, nv)
        R = dense.kron(p2[ti]) @ Dm @ dense.kron(p1[ti])
        out = np.zeros(4, complex)
        for iop in range(4):
            G = _tables.G5 @ np.array([[_tables.rhoj(iop + 1, idir, r, c) for c in range(4)]
                                       for r in range(4)])
            ip = _tables.compute_ip(idir, iop)
            out[iop] = np.trace(dense.gam(G, nv) @ dense.kron(p12x5[tf, ip]) @ T @ g5 @ R @ T.conj().T)
        return out

    @staticmethod
    def pipi(tau_b, diag_tf, diag_ti, m1, m2, m3, m4, tf, ti):
        """Unweighted [square, doubleTriangle, doubleTriangleHor, star(t1,t2), fish(t1,t2)]."""
        nv = tau_b.shape[-1]
        T = dense.tau(tau_b)
        Th = T.conj().T
        g5 = dense.gam(_tables.G5, nv)
        K = dense.kron
        Dtf, Dti = dense.tau(diag_tf), dense.tau(diag_ti)
        Q = K(m3[tf]) @ Dtf @ K(m4[tf])
        sq = np.trace(Q @ g5 @ T @ g5 @ K(m2[ti]) @ Dti @ K(m1[ti]) @ Th)
        dt = np.trace(Q @ g5 @ T @ g5 @ K(m1[ti]) @ Dti @ K(m2[ti]) @ Th)
        W = dense.w_hor(tau_b)
        hor = np.trace(K(m3[tf]) @ T @ K(m1[ti]) @ W @ K(m4[tf]) @ T @ K(m2[ti]) @ Th)
        st = (np.trace(K(m4[tf]) @ T @ K(m1[ti]) @ Th), np.trace(K(m3[tf]) @ T @ K(m2[ti]) @ Th))
        fi = (np.trace(K(m3[tf]) @ T @ K(m1[ti]) @ Th), np.trace(K(m4[tf]) @ T @ K(m2[ti]) @ Th))
        return sq, dt, hor, st, fi

    @staticmethod
    def loop(diag_b, m):
        return np.trace(dense.tau(diag_b) @ dense.kron(m))

    @staticmethod
    def sigma(tau_b, x2, tf, ti):
        """[4] indexed lp*2+rp: Tr[(X_lp(tf) tau)^dag g5 (tau X_rp(ti)) g5]."""
        nv = tau_b.shape[-1]
        T = dense.tau(tau_b)
        g5 = dense.gam(_tables.G5, nv)
        out = np.zeros(4, complex)
        for lp in range(2):
            for rp in range(2):
                B = dense.kron(x2[tf, lp]) @ T
                A = T @ dense.kron(x2[ti, rp])
                out[lp * 2 + rp] = np.trace(B.conj().T @ g5 @ A @ g5)
        return out

    @staticmethod
    def fish_hor(diag_b, p1, p2, t):
        Dm = dense.tau(diag_b)
        return np.trace(dense.kron(p2[t]) @ Dm @ dense.kron(p1[t]) @ Dm.conj().T)

    @staticmethod
    def triangle(tau_b, diag_b, p1, p2, x2, tf, ti):
        nv = tau_b.shape[-1]
        T = dense.tau(tau_b)
        g5 = dense.gam(_tables.G5, nv)
        R = dense.kron(p2[ti]) @ dense.tau(diag_b) @ dense.kron(p1[ti])
        return np.array([np.trace(g5 @ dense.kron(x2[tf, ip]) @ T @ g5 @ R @ T.conj().T)
                         for ip in range(2)])
