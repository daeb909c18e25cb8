"""Torch-facing wrapper of the batched engine (lc_engine_rho_pipi).

tau: DeviceArray complex128 [nt][nt][4][4][nv][nv]; phi: DeviceArray complex128 [np][nt][nv][nv].
Returns (raw, corr, avg) DeviceArrays:
  raw  [np][nsrc][nt][10]  {rho_0..2, square, dT, dTHor, star1, star2, fish1, fish2} traces
  corr [np][nsrc][nt][8]   {rho_0..2, square, dT, dTHor, star, fish}
  avg  [np][8][nt]         source average C(dt)
"""
import ctypes

from ._capi import check, lib, register

CORR_NAMES = ["rho_x", "rho_y", "rho_z", "square", "doubleTriangle", "doubleTriangleHor", "star", "fish"]
N_CORR = len(CORR_NAMES)

_VP, _I, _LL = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong
register("lc_engine_rho_pipi", [_VP, _VP, _I, _I, _I, _I, _I, _I, _LL, _VP, _VP, _VP])
register("lc_engine_rho_pipi_mixed", [_VP, _VP, _I, _I, _I, _I, _I, _I, _LL, _VP, _VP, _VP])
register("lc_engine_pipi_sets", [_VP, _VP, _I, _I, _I, _I, _I, _I, _LL, _VP, _VP, _VP])
register("lc_engine_pipi_sets_mixed", [_VP, _VP, _I, _I, _I, _I, _I, _I, _LL, _VP, _VP, _VP])
register("lc_engine_sigma_set", [_VP, _VP, _I, _I, _I, _I, _I, _I, _I, _LL, _VP, _VP])
register("lc_engine_sigma_set_mixed", [_VP, _VP, _I, _I, _I, _I, _I, _I, _I, _LL, _VP, _VP])
register("lc_selftest_mfma_bf16", [_VP, _VP, _VP])
register("lc_build_elementals", [_VP, _VP, _I, _I, _LL, _I, _VP])
register("lc_covariant_derivative", [_VP, _VP, _I, _I, _I, _I, _I, _VP])
register("lc_covariant_derivative_scaled", [_VP, _VP, _I, _I, _I, _I, _I, ctypes.c_double, _VP])
register("lc_build_elementals_types", [_VP, _VP, _VP, _I, _I, _LL, _I, _VP])
register("lc_profile", [_I])
register("lc_set_tile_params", [_I])
register("lc_set_dthor_dagger", [_I])
register("lc_engine_graph", [_I])
register("lc_engine_graph_report", [_VP])
register("lc_autotune", [_I])
register("lc_autotune_report", [_VP])
register("lc_profile_read", [_VP])
register("lc_set_mixed_form", [_I])

MIXED_EXACT, MIXED_BF16X3 = 0, 1


def set_mixed_form(form):
    """Form of the mixed-precision engines: MIXED_EXACT (default; int8 digits with per-line
    exponents, exact integer accumulation, FP64 products and traces) or MIXED_BF16X3 (per-slice
    rescaled FP32 storage, BF16x3 split products with FP32 accumulation)."""
    check(lib().lc_set_mixed_form(int(form)))


def rho_pipi(tau, phi, src0=0, src_step=1, nsrc=None, max_scratch_bytes=0, out=None, mixed=False):
    """mixed=False: plain FP64 on the FP64 matrix cores; mixed=True: the mixed form selected by
    set_mixed_form (default: int8 digits with exact integer accumulation, FP64 traces and sums)."""
    import numpy as np
    from .devmem import DeviceArray
    if not (isinstance(tau, DeviceArray) and isinstance(phi, DeviceArray)):
        raise TypeError("engine operands must be DeviceArrays")
    if tau.dtype != np.complex128 or phi.dtype != np.complex128:
        raise TypeError("engine operands are complex128")
    nt, nv = tau.shape[0], tau.shape[-1]
    if tuple(tau.shape) != (nt, nt, 4, 4, nv, nv):
        raise ValueError("tau must be [nt][nt][4][4][nv][nv]")
    np_ = phi.shape[0]
    if tuple(phi.shape) != (np_, nt, nv, nv):
        raise ValueError("phi must be [np][nt][nv][nv]")
    if nsrc is None:
        nsrc = (nt - src0 + src_step - 1) // src_step
    if out is None:
        raw = DeviceArray((np_, nsrc, nt, 10), np.complex128)
        corr = DeviceArray((np_, nsrc, nt, N_CORR), np.complex128)
        avg = DeviceArray((np_, N_CORR, nt), np.complex128)
    else:
        raw, corr, avg = out
        if (raw.shape != (np_, nsrc, nt, 10) or corr.shape != (np_, nsrc, nt, N_CORR)
                or avg.shape != (np_, N_CORR, nt)):
            raise ValueError("output buffers have the wrong shape")
    fn = lib().lc_engine_rho_pipi_mixed if mixed else lib().lc_engine_rho_pipi
    check(fn(tau.ptr, phi.ptr, nt, nv, np_, src0, src_step, nsrc, max_scratch_bytes,
             raw.ptr, corr.ptr, avg.ptr))
    return raw, corr, avg


def pipi_sets(tau, elem, src0=0, src_step=1, nsrc=None, max_scratch_bytes=0, out=None, mixed=False):
    """Pi-pi over explicit momentum sets (lc_engine_pipi_sets[_mixed]).
    tau: DeviceArray complex128 [nt][nt][4][4][nv][nv]; elem: complex128 [nset][4][nt][nv][nv],
    the elementals P1, P2 (source) and P3, P4 (sink) of each set, the momP1..momP4 of
    contractPiPi (lib/contract.cpp:748). Returns (raw, corr, avg) as rho_pipi with np = nset."""
    import numpy as np
    from .devmem import DeviceArray
    if not (isinstance(tau, DeviceArray) and isinstance(elem, DeviceArray)):
        raise TypeError("engine operands must be DeviceArrays")
    if tau.dtype != np.complex128 or elem.dtype != np.complex128:
        raise TypeError("engine operands are complex128")
    nt, nv = tau.shape[0], tau.shape[-1]
    if tuple(tau.shape) != (nt, nt, 4, 4, nv, nv):
        raise ValueError("tau must be [nt][nt][4][4][nv][nv]")
    ns_ = elem.shape[0]
    if tuple(elem.shape) != (ns_, 4, nt, nv, nv):
        raise ValueError("elem must be [nset][4][nt][nv][nv]")
    if nsrc is None:
        nsrc = (nt - src0 + src_step - 1) // src_step
    if out is None:
        raw = DeviceArray((ns_, nsrc, nt, 10), np.complex128)
        corr = DeviceArray((ns_, nsrc, nt, N_CORR), np.complex128)
        avg = DeviceArray((ns_, N_CORR, nt), np.complex128)
    else:
        raw, corr, avg = out
        if (raw.shape != (ns_, nsrc, nt, 10) or corr.shape != (ns_, nsrc, nt, N_CORR)
                or avg.shape != (ns_, N_CORR, nt)):
            raise ValueError("output buffers have the wrong shape")
    fn = lib().lc_engine_pipi_sets_mixed if mixed else lib().lc_engine_pipi_sets
    check(fn(tau.ptr, elem.ptr, nt, nv, ns_, src0, src_step, nsrc, max_scratch_bytes,
             raw.ptr, corr.ptr, avg.ptr))
    return raw, corr, avg


SIGMA_RAW_NAMES = ["sigma_00", "sigma_01", "sigma_10", "sigma_11", "triangle_0", "triangle_1",
                   "rhopi_0", "rhopi_1", "rhopi_2", "rhopi_3"]
EL_TYPES = 11  # 0-1 sigma X_l, 2-3 triangle X_ip, 4-8 rho-pi P1P2 types, 9 P1, 10 P2


def sigma_set(tau, el, idir=0, src0=0, src_step=1, nsrc=None, max_scratch_bytes=0, out=None, mixed=False):
    """Batched sigma / sigma-pi / rho-pi set (lc_engine_sigma_set[_mixed]).
    tau: DeviceArray complex128 [nt][nt][4][4][nv][nv]; el: complex128 [np][nt][11][nv][nv].
    Returns (raw [np][nsrc][nt][10], loops [np][nt][3]) DeviceArrays."""
    import numpy as np
    from .devmem import DeviceArray
    if not (isinstance(tau, DeviceArray) and isinstance(el, DeviceArray)):
        raise TypeError("engine operands must be DeviceArrays")
    if tau.dtype != np.complex128 or el.dtype != np.complex128:
        raise TypeError("engine operands are complex128")
    nt, nv = tau.shape[0], tau.shape[-1]
    if tuple(tau.shape) != (nt, nt, 4, 4, nv, nv):
        raise ValueError("tau must be [nt][nt][4][4][nv][nv]")
    np_ = el.shape[0]
    if tuple(el.shape) != (np_, nt, EL_TYPES, nv, nv):
        raise ValueError("el must be [np][nt][11][nv][nv]")
    if nsrc is None:
        nsrc = (nt - src0 + src_step - 1) // src_step
    if out is None:
        raw = DeviceArray((np_, nsrc, nt, 10), np.complex128)
        loops = DeviceArray((np_, nt, 3), np.complex128)
    else:
        raw, loops = out
        if raw.shape != (np_, nsrc, nt, 10) or loops.shape != (np_, nt, 3):
            raise ValueError("output buffers have the wrong shape")
    fn = lib().lc_engine_sigma_set_mixed if mixed else lib().lc_engine_sigma_set
    check(fn(tau.ptr, el.ptr, nt, nv, np_, idir, src0, src_step, nsrc, max_scratch_bytes, raw.ptr, loops.ptr))
    return raw, loops


def profile(enable=True):
    check(lib().lc_profile(1 if enable else 0))


def profile_read():
    import numpy as np
    buf = np.zeros(6)
    check(lib().lc_profile_read(buf.ctypes.data))
    return dict(gemm_ms=buf[0], gemm_launches=int(buf[1]), gemm_flops=buf[2],
                trace_ms=buf[3], trace_launches=int(buf[4]), trace_flops=buf[5])


def graph(enable=True):
    """hipGraph replay of the engines: the first call with given arguments is recorded, later calls
    with the same arguments replay it (lc_engine_graph); False destroys the graphs."""
    check(lib().lc_engine_graph(1 if enable else 0))


def graph_report():
    """What the last engine call did: "eager", "recorded" or "replayed"."""
    v = ctypes.c_int(0)
    check(lib().lc_engine_graph_report(ctypes.byref(v)))
    return ("eager", "recorded", "replayed")[v.value]


def set_dthor_dagger(on=True):
    """doubleTriangleHor with the conjugate transpose tau^dag instead of the reference's element-wise
    conjugate (lib/calculate_pipi_doubleTriangleHor_core.h:126); applies to the pi-pi row launcher,
    calculatePiPi and the rho + pi-pi engines. Off by default."""
    check(lib().lc_set_dthor_dagger(1 if on else 0))


def set_tile_params(tiles_per_wg=0):
    """Output tiles (64 x 64) per workgroup of the mixed GEMM: 1, 2 or 4; 0 = built-in rule."""
    check(lib().lc_set_tile_params(int(tiles_per_wg)))


def autotune(enable=True, forget=False):
    """Turn the in-project autotuner on or off (forget=True also drops its cached choices)."""
    check(lib().lc_autotune(-1 if forget else (1 if enable else 0)))
    if forget and enable:
        check(lib().lc_autotune(1))


def autotune_report():
    """Tile shape of the last engine call: tiles per workgroup (0: rule), how it was chosen
    ("fixed", "cached", "tuned") and the candidate timings in ms (4, 2, 1 tiles per workgroup)."""
    import numpy as np
    buf = np.zeros(5)
    check(lib().lc_autotune_report(buf.ctypes.data))
    return dict(tiles_per_wg=int(buf[0]), how=("fixed", "cached", "tuned")[int(buf[1])],
                candidate_ms={4: buf[2], 2: buf[3], 1: buf[4]})


def selftest_mfma_bf16(a, b):
    """One v_mfma_f32_16x16x32_bf16 on A (16x32), B (32x16) exact in BF16: (D, raw [64][4])."""
    import numpy as np
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    d = np.zeros(512, np.float32)
    check(lib().lc_selftest_mfma_bf16(a.ctypes.data, b.ctypes.data, d.ctypes.data))
    return d[:256].reshape(16, 16), d[256:].reshape(64, 4)


def build_elementals(evec, phase, out=None):
    """Plain momentum elementals on the GPU (lc_build_elementals): evec [nt][nv][nsite] and phase
    [np][nsite] DeviceArrays (complex128) -> DeviceArray [np][nt][nv][nv] with
    Phi[p][t][r][c] = sum_k conj(evec[t][r][k]) phase[p][k] evec[t][c][k]
    (the drivers' host loop, tests/compute_rho.cpp:116-129)."""
    import numpy as np
    from .devmem import DeviceArray
    if not (isinstance(evec, DeviceArray) and isinstance(phase, DeviceArray)):
        raise TypeError("operands must be DeviceArrays")
    if evec.dtype != np.complex128 or phase.dtype != np.complex128:
        raise TypeError("operands are complex128")
    nt, nv, nsite = evec.shape
    np_ = phase.shape[0]
    if tuple(phase.shape) != (np_, nsite):
        raise ValueError("phase must be [np][nsite] with the eigenvectors' nsite")
    if out is None:
        out = DeviceArray((np_, nt, nv, nv), np.complex128)
    check(lib().lc_build_elementals(evec.ptr, phase.ptr, nt, nv, nsite, np_, out.ptr))
    return out


def covariant_derivative(links, evec, nx, ny, nz, out=None, c=0.5):
    """Covariant derivatives of the eigenvectors on the GPU (lc_covariant_derivative_scaled):
    links [nt][3][nx*ny*nz][3][3], evec [nt][nv][3*nx*ny*nz] DeviceArrays (complex128) ->
    DeviceArray [3][nt][nv][3*nx*ny*nz], (nabla_j v)(x) = c [U_j(x) v(x+j) - U_j(x-j)^dag v(x-j)]; c is
    the stencil normalisation of the drivers' external covderiv (not in the reference; 1/2 here by
    default, the central difference)."""
    import numpy as np
    from .devmem import DeviceArray
    if not (isinstance(links, DeviceArray) and isinstance(evec, DeviceArray)):
        raise TypeError("operands must be DeviceArrays")
    if links.dtype != np.complex128 or evec.dtype != np.complex128:
        raise TypeError("operands are complex128")
    nt, nv, nk = evec.shape
    ns = nx * ny * nz
    if nk != 3 * ns:
        raise ValueError("evec must be [nt][nv][3*nx*ny*nz]")
    if tuple(links.shape) != (nt, 3, ns, 3, 3):
        raise ValueError("links must be [nt][3][nx*ny*nz][3][3]")
    if out is None:
        out = DeviceArray((3, nt, nv, nk), np.complex128)
    elif tuple(out.shape) != (3, nt, nv, nk):
        raise ValueError("out must be [3][nt][nv][3*nx*ny*nz]")
    check(lib().lc_covariant_derivative_scaled(links.ptr, evec.ptr, nt, nv, nx, ny, nz, float(c), out.ptr))
    return out


def build_elementals_types(evec, dvec, phase, out=None):
    """The drivers' five elemental types (lc_build_elementals_types): evec [nt][nv][nsite], dvec
    [3][nt][nv][nsite] (covariant_derivative), phase [np][nsite] -> DeviceArray
    [np][5][nt][nv][nv]: plain, smeared sum_j dV_j^dag ph dV_j, and V^dag ph dV_j - dV_j^dag ph V
    (tests/compute_rho.cpp:116-229)."""
    import numpy as np
    from .devmem import DeviceArray
    for x in (evec, dvec, phase):
        if not isinstance(x, DeviceArray):
            raise TypeError("operands must be DeviceArrays")
        if x.dtype != np.complex128:
            raise TypeError("operands are complex128")
    nt, nv, nsite = evec.shape
    np_ = phase.shape[0]
    if tuple(dvec.shape) != (3, nt, nv, nsite):
        raise ValueError("dvec must be [3][nt][nv][nsite]")
    if tuple(phase.shape) != (np_, nsite):
        raise ValueError("phase must be [np][nsite] with the eigenvectors' nsite")
    if out is None:
        out = DeviceArray((np_, 5, nt, nv, nv), np.complex128)
    elif tuple(out.shape) != (np_, 5, nt, nv, nv):
        raise ValueError("out must be [np][5][nt][nv][nv]")
    check(lib().lc_build_elementals_types(evec.ptr, dvec.ptr, phase.ptr, nt, nv, nsite, np_, out.ptr))
    return out


def errors_per_t(c_test, c_ref, corr_ref=None, src0=0, src_step=1):
    """Per source-time separation t: max over (momentum, diagram) of |dC(t)|/|C(t)| and of
    |dC(t)|/max_t|C(t)|, and the fraction of all C(t) within 1e-6 relative; the same element-wise
    bound over the significant C(t) (|C(t)| >= 1e-3 max_t |C(t)| of their correlator), with the
    fraction of C(t) that are significant and of those within 1e-6. With corr_ref (the
    reference per-source correlators [np][nsrc][nt][ncorr]) also |dC(t)| / S(t), where S(t) is the
    source average of |c(t0 + t, t0)|: the scale of the summed terms, which a C(t) that happens to
    cancel between sources does not have."""
    import numpy as np
    d = np.abs(c_test - c_ref)          # [np][ncorr][nt]
    a = np.abs(c_ref)
    rel = d / np.where(a > 0, a, np.inf)
    norm = d / a.max(axis=2, keepdims=True)
    sig = a >= 1e-3 * a.max(axis=2, keepdims=True)  # the C(t) of at least 1e-3 of their correlator's max
    rs = np.where(sig, rel, 0.0)
    out = dict(max_rel=[float('%.3g' % v) for v in rel.max(axis=(0, 1))],
               max_normwise=[float('%.3g' % v) for v in norm.max(axis=(0, 1))],
               frac_rel_le_1e6=float((rel <= 1e-6).mean()),
               max_rel_significant=[float('%.3g' % v) for v in rs.max(axis=(0, 1))],
               frac_significant_rel_le_1e6=float((rel[sig] <= 1e-6).mean()) if sig.any() else 1.0,
               significant_fraction=float(sig.mean()))
    if corr_ref is not None:
        np_, nsrc, nt, nc = corr_ref.shape
        scale = np.zeros((np_, nc, nt))
        for k in range(nsrc):
            t0 = src0 + k * src_step
            for dt in range(nt):
                scale[:, :, dt] += np.abs(corr_ref[:, k, (t0 + dt) % nt, :])
        scale /= nsrc
        rs = d / np.where(scale > 0, scale, np.inf)
        out["max_rel_source_scale"] = [float('%.3g' % v) for v in rs.max(axis=(0, 1))]
        out["max_rel_source_scale_all"] = float(rs.max())
    return out


def rel_errors(c_test, c_ref):
    """Per diagram: max over (momentum, t) of |dC|/|C|, of |dC|/max_t|C| (normwise) and of |dC|/|C|
    over the significant C(t) (|C(t)| >= 1e-3 max_t |C(t)|)."""
    import numpy as np
    d = np.abs(c_test - c_ref)          # [np][ncorr][nt]
    a = np.abs(c_ref)
    rel = d / np.where(a > 0, a, np.inf)
    norm = d / a.max(axis=2, keepdims=True)
    sig = a >= 1e-3 * a.max(axis=2, keepdims=True)
    rs = np.where(sig, rel, 0.0)
    return {name: dict(max_rel=float(rel[:, k].max()), max_normwise=float(norm[:, k].max()),
                       max_rel_significant=float(rs[:, k].max()))
            for k, name in enumerate(CORR_NAMES)}
