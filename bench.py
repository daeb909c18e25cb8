"""Benchmark: correlators/s of the rho + pi-pi set at D = 256 on one MI355X.

Workload (one step = one full pass): Nv = 64 eigenvectors x 4 spins (D = 256), Lt = 48 time
slices, all 48 source times, 19 momenta; per (sink tf, source t0, momentum p) 8 correlators:
rho with Gamma = gamma_x, gamma_y, gamma_z, and the pi-pi square, doubleTriangle,
doubleTriangleHor, star and fish diagrams of back-to-back pions (Phi(p), Phi(p)^dag). Inputs are
synthetic and seeded: tau(t,t0) i.i.d. CN(0,1) exp(-0.3 |t-t0|)/sqrt(D), Phi(0,t) Hermitian,
Phi(p!=0,t) i.i.d. CN(0,1)/sqrt(Nv); they are resident in HBM before the timed region.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-rows R]

Prints one JSON line. The CPU baseline is the oracle (the reference-structured restatement,
one table row at a time, OpenMP) on a bounded sample of sink rows of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "correlators/s at D=256, mixed vs FP64 on one GPU at ≤1e-6 rel err; %MFMA peak"
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16": 2500.0, "int8": 5000.0}  # dense; int8 in TOP/s
# The exact mixed form computes a complex multiply-add (8 algorithmic flops) as 10 digit pairs x 4
# real int8 products = 40 int8 multiply-adds (80 int8 ops): 10 int8 ops per algorithmic flop. The
# roofline's peak is the published int8 dense peak; the issue utilisation of the int8 matrix cores
# (algorithmic rate x 10 / int8 peak) is reported beside it.
EXACT_OPS_PER_FLOP = 10.0
HBM_PEAK_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def workload(nv=64, nt=48, nmom=19, seed=3, mass=0.3):
    import numpy as np
    from contractions_code_scattering_amd import inputs as I
    g = I.rng(seed)
    tau = I.perambulator(g, nt, nv, mass)
    phi = [I.hermitian_elemental(g, nt, nv)] + [I.random_elemental(g, nt, nv) for _ in range(nmom - 1)]
    return tau, np.stack(phi)


def cpu_baseline(tau, phi, rows):
    """Oracle on `rows` sink rows of momentum 0: rho op pair (0,0) + the I=0 pi-pi set.
    Returns the baseline record and the oracle's correlators [rows][nt][8] for a parity check."""
    import numpy as np
    import oracle as O
    from contractions_code_scattering_amd import inputs as I
    nt = tau.shape[0]
    P = phi[0]
    Pd = I.dagger(P)
    mom5 = np.ascontiguousarray(np.stack([P] * 5, axis=1))
    diag = O.diag_of(tau)
    ref = np.zeros((rows, nt, 8), complex)
    t = time.perf_counter()
    for tf in range(rows):
        r = O.rho_row_range(tau[tf], mom5, tf, 0, 48)
        q = O.pipi_row(0, tau[tf], diag, P, Pd, Pd, P, tf)
        ref[tf, :, 0:3] = r[:, 0, :]
        ref[tf, :, 3] = -3.0 * q[:, 0]
        ref[tf, :, 4] = -3.0 * q[:, 1]
        ref[tf, :, 5] = -0.6 * q[:, 4]
        ref[tf, :, 6] = q[:, 2]
        ref[tf, :, 7] = q[:, 3]
    dt = time.perf_counter() - t
    ncorr = rows * nt * 8
    rec = dict(value=ncorr / dt, unit="correlators/s", cores=O.num_threads(), kind="port",
               sample="%d sink rows x %d sources x 8 correlators of momentum 0 (rho op pair (0,0) "
                      "+ pi-pi I=0 set) at Nv=64, Lt=48, FP64, %.1f s" % (rows, nt, dt),
               cpu_model=cpu_model(), host_cpus=os.cpu_count())
    return rec, ref


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def traffic_from_profiles(kernel, nmom):
    """HBM bytes per launch of the kernel whose name contains `kernel`, from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (calibrated against a 1 GiB copy); None when no summary exists
    for this workload."""
    path = os.path.join(ROOT, "profiles", "traffic_rho_pipi_D256_Lt48_%dmom.json" % nmom)
    try:
        with open(path) as f:
            ks = json.load(f)["kernels"]
        for name, k in ks.items():
            if kernel in name:
                return round(k["bytes_per_launch"])
        return None
    except (OSError, ValueError, KeyError):
        return None


def _two_forms(label, call, ncorr, collect):
    """FP64 then mixed: call(mixed) once untimed (the mixed form tunes its tile shape), once timed.
    Returns ({fp64, mixed, mixed_vs_fp64, unit}, {False: collect(outputs), True: ...})."""
    from contractions_code_scattering_amd import engine
    from contractions_code_scattering_amd.devmem import synchronize
    res, outs = {}, {}
    for mixed in (False, True):
        call(mixed)
        synchronize()
        tune = engine.autotune_report()
        engine.profile(True)
        t = time.perf_counter()
        out = call(mixed)
        synchronize()
        dt = time.perf_counter() - t
        p = engine.profile_read()
        engine.profile(False)
        outs[mixed] = collect(out)
        key = "mixed" if mixed else "fp64"
        res[key] = dict(value=round(ncorr / dt, 1), ms=round(dt * 1e3, 2),
                        gemm_tflops=round(p["gemm_flops"] / (p["gemm_ms"] * 1e-3) / 1e12, 2) if p["gemm_ms"] else None)
        if mixed and tune.get("how") in ("tuned", "cached"):
            res[key]["tiles_per_wg"] = tune["tiles_per_wg"]
        log("%s %s: %.1f ms" % (label, key, dt * 1e3))
    res["mixed_vs_fp64"] = round(res["mixed"]["value"] / res["fp64"]["value"], 3)
    res["unit"] = "correlators/s"
    return res, outs


def d512_measurement(nmom, nv=128, nt=64, seed=5, mass=0.25):
    """rho + pi-pi at the largest size: Nv = 128 (D = 512), Lt = 64, all 64 sources, `nmom` momenta
    (the first Hermitian), perambulator decay 0.25; both forms, one timed pass each after one
    warm-up pass. At Nv = 128 the mixed form runs its four-level chain (no fused sandwiches).
    On the same resident perambulator, also a moving frame's pi-pi over `nmom` explicit momentum
    sets and the sigma / sigma-pi / rho-pi set for `nmom` momentum sets (pipi_sets, sigma_set)."""
    import numpy as np
    from contractions_code_scattering_amd import engine
    from contractions_code_scattering_amd.devmem import DeviceArray, synchronize, to_device
    tau_h, phi_h = workload(nv, nt, nmom, seed, mass)
    tau, phi = to_device(tau_h), to_device(phi_h)
    del tau_h
    log("d512 inputs resident: tau %.2f GB" % (tau.nbytes / 1e9))
    ncorr = nmom * nt * nt * 8
    res, avg = {}, {}
    for mixed in (False, True):
        out = (DeviceArray((nmom, nt, nt, 10), np.complex128), DeviceArray((nmom, nt, nt, 8), np.complex128),
               DeviceArray((nmom, 8, nt), np.complex128))
        engine.rho_pipi(tau, phi, out=out, mixed=mixed)  # warm-up (the mixed form tunes its tile shape)
        synchronize()
        tune = engine.autotune_report()
        engine.profile(True)
        t = time.perf_counter()
        engine.rho_pipi(tau, phi, out=out, mixed=mixed)
        synchronize()
        dt = time.perf_counter() - t
        p = engine.profile_read()
        engine.profile(False)
        avg[mixed] = out[2].to_numpy()
        key = "mixed" if mixed else "fp64"
        res[key] = dict(value=round(ncorr / dt, 1), ms=round(dt * 1e3, 2),
                        gemm_tflops=round(p["gemm_flops"] / (p["gemm_ms"] * 1e-3) / 1e12, 2) if p["gemm_ms"] else None)
        if mixed and tune.get("how") in ("tuned", "cached"):
            res[key]["tiles_per_wg"] = tune["tiles_per_wg"]
            res[key]["tile_candidates_ms"] = {k: round(v, 2) for k, v in tune["candidate_ms"].items()}
        log("d512 %s: %.1f ms" % (key, dt * 1e3))
    errs = engine.rel_errors(avg[True], avg[False])
    res["mixed_vs_fp64"] = round(res["mixed"]["value"] / res["fp64"]["value"], 3)
    res["max_normwise_err"] = max(e["max_normwise"] for e in errs.values())
    res["max_rel_err"] = max(e["max_rel"] for e in errs.values())
    res["unit"] = "correlators/s"
    res["config"] = dict(workload="rho(3 pol)+pipi(5 diagrams)", nv=nv, D=4 * nv, lt=nt, sources=nt, momenta=nmom,
                         mass=mass, seed=seed, correlators_per_pass=ncorr)
    del phi
    from contractions_code_scattering_amd import inputs as I
    g = I.rng(seed + 100)
    # a moving frame (total momentum (0,0,1)): its pion pairs are not back-to-back, so each momentum
    # set carries four independent elementals P1..P4 (random, distributed as the workload's Phi)
    el = to_device(np.stack([np.stack([I.random_elemental(g, nt, nv) for _ in range(4)]) for _ in range(nmom)]))
    r, o = _two_forms("d512 pipi sets", lambda mixed: engine.pipi_sets(tau, el, mixed=mixed), ncorr,
                      lambda out: out[2].to_numpy())
    e = engine.rel_errors(o[True], o[False])
    r["max_normwise_err"] = max(v["max_normwise"] for v in e.values())
    r["config"] = dict(workload="pipi(5 diagrams)+rho slots over explicit momentum sets P1..P4", sets=nmom,
                       correlators_per_pass=ncorr)
    res["pipi_sets"] = r
    del el
    # the sigma / sigma-pi / rho-pi set on the same perambulator, all 64 sources, idir = 2
    el = to_device(np.stack([I.cn(g, (nt, 11, nv, nv)) / np.sqrt(nv) for _ in range(nmom)]))
    nc = nmom * nt * nt * 10
    r, o = _two_forms("d512 sigma set", lambda mixed: engine.sigma_set(tau, el, idir=2, mixed=mixed), nc,
                      lambda out: out[0].to_numpy())
    d = np.abs(o[True] - o[False]).reshape(-1, 10)
    a = np.abs(o[False]).reshape(-1, 10)
    r["max_normwise_err"] = float((d.max(axis=0) / a.max(axis=0)).max())
    r["config"] = dict(workload="sigma one_fish (4) + sigma-pi triangle (2) + rho-pi (4) per (tf, t0, p)",
                       momentum_sets=nmom, correlators_per_pass=nc)
    res["sigma_set"] = r
    del el
    return res


def elementals_measurement(L=24, nv=64, nt=48, nmom=19, seed=6):
    """Plain momentum elementals Phi(p, t) = V_t^dag e^{-ip.x} V_t on the GPU for an L^3 lattice with 3
    colours (K = 3 L^3), Nv eigenvectors per time slice (random), the 19 momenta |p|^2 <= 2; one timed
    call after a warm-up call."""
    import numpy as np
    from contractions_code_scattering_amd import engine
    from contractions_code_scattering_amd import inputs as I
    from contractions_code_scattering_amd.devmem import synchronize, to_device
    g = I.rng(seed)
    nsite = 3 * L ** 3
    V = to_device(I.cn(g, (nt, nv, nsite)))
    moms = [(0, 0, 0)] + [p for p in [(1, 0, 0), (0, 1, 0), (0, 0, 1), (-1, 0, 0), (0, -1, 0), (0, 0, -1),
                                      (1, 1, 0), (1, -1, 0), (-1, 1, 0), (-1, -1, 0), (1, 0, 1), (1, 0, -1),
                                      (-1, 0, 1), (-1, 0, -1), (0, 1, 1), (0, 1, -1), (0, -1, 1), (0, -1, -1)]]
    moms = moms[:nmom]
    x = np.indices((L, L, L)).reshape(3, -1)
    ph = np.stack([np.repeat(np.exp(-2j * np.pi * (np.array(p) @ x) / L), 3) for p in moms])
    phd = to_device(np.ascontiguousarray(ph))
    out = engine.build_elementals(V, phd)
    synchronize()
    t = time.perf_counter()
    engine.build_elementals(V, phd, out=out)
    synchronize()
    dt = time.perf_counter() - t
    flops = 8.0 * nmom * nt * nv * nv * nsite
    log("elementals: %.2f ms" % (dt * 1e3))
    res = dict(ms=round(dt * 1e3, 3), tflops=round(flops / dt / 1e12, 2),
               frac_f64_peak=round(flops / dt / 1e12 / PEAK_TFLOPS["f64"], 3),
               config=dict(L=L, colours=3, nsite=nsite, nv=nv, lt=nt, momenta=nmom, dtype="f64"))
    # the derivative types: covariant derivatives (links random complex 3x3; the stencil's cost does
    # not depend on them being SU(3)), then the five types per momentum (10 bilinears each)
    U = to_device(I.cn(g, (nt, 3, L ** 3, 3, 3)))
    dV = engine.covariant_derivative(U, V, L, L, L)
    synchronize()
    t = time.perf_counter()
    engine.covariant_derivative(U, V, L, L, L, out=dV)
    synchronize()
    dt_d = time.perf_counter() - t
    vbytes = nt * nv * nsite * 16
    dbytes = 4 * vbytes + U.nbytes  # V read once, U once, 3 V written
    T = engine.build_elementals_types(V, dV, phd)
    synchronize()
    t = time.perf_counter()
    engine.build_elementals_types(V, dV, phd, out=T)
    synchronize()
    dt_t = time.perf_counter() - t
    log("covariant derivative: %.2f ms; five types: %.2f ms" % (dt_d * 1e3, dt_t * 1e3))
    res["covariant_derivative"] = dict(ms=round(dt_d * 1e3, 3), gbs=round(dbytes / dt_d / 1e9, 1),
                                       frac_hbm_peak=round(dbytes / dt_d / 1e9 / HBM_PEAK_GBS, 3),
                                       algorithmic_bytes=dbytes)
    res["types"] = dict(ms=round(dt_t * 1e3, 3), tflops=round(10 * flops / dt_t / 1e12, 2),
                        frac_f64_peak=round(10 * flops / dt_t / 1e12 / PEAK_TFLOPS["f64"], 3),
                        bilinears_per_momentum=10)
    return res


def sigma_set_measurement(nv=64, nt=64, nmom=27, src_step=4, seed=4, bf16x3=False, only=None):
    """The batched sigma / sigma-pi / rho-pi set on its own workload, one timed pass per form after
    one warm-up pass: Nv = 64 (D = 256), Lt = 64, sources every 4th time slice (16), 27 momentum
    sets of 11 elemental types each, rho-pi polarisation idir = 2. Forms: FP64, the selected mixed
    form ("mixed"), and with the exact form selected also BF16x3 for comparison; only = "fp64" or
    "mixed": that form alone (profiling)."""
    import numpy as np
    from contractions_code_scattering_amd import engine
    from contractions_code_scattering_amd import inputs as I
    from contractions_code_scattering_amd.devmem import synchronize, to_device
    g = I.rng(seed)
    tau = to_device(I.perambulator(g, nt, nv, 0.3))
    el = to_device(np.stack([I.cn(g, (nt, 11, nv, nv)) / np.sqrt(nv) for _ in range(nmom)]))
    nsrc = nt // src_step
    ncorr = nmom * nsrc * nt * 10
    res = {}
    outs = {}
    sel = engine.MIXED_BF16X3 if bf16x3 else engine.MIXED_EXACT
    forms = [("fp64", False, sel), ("mixed", True, sel)]
    if not bf16x3:
        forms.append(("bf16x3", True, engine.MIXED_BF16X3))
    if only:
        forms = [f for f in forms if f[0] == only]
    for key, mixed, form in forms:
        engine.set_mixed_form(form)
        engine.sigma_set(tau, el, idir=2, src_step=src_step, mixed=mixed)  # warm-up (BF16x3: tunes the tile shape)
        synchronize()
        tune = engine.autotune_report()
        engine.profile(True)
        t = time.perf_counter()
        raw, loops = engine.sigma_set(tau, el, idir=2, src_step=src_step, mixed=mixed)
        synchronize()
        dt = time.perf_counter() - t
        p = engine.profile_read()
        engine.profile(False)
        outs[key] = raw.to_numpy()
        res[key] = dict(value=round(ncorr / dt, 1), ms=round(dt * 1e3, 2),
                        gemm_tflops=round(p["gemm_flops"] / (p["gemm_ms"] * 1e-3) / 1e12, 2) if p["gemm_ms"] else None)
        if mixed:
            res[key]["form"] = "bf16x3" if form == engine.MIXED_BF16X3 else "exact (int8 digits)"
        if mixed and form == engine.MIXED_BF16X3 and tune.get("how") in ("tuned", "cached"):
            res[key]["tiles_per_wg"] = tune["tiles_per_wg"]
            res[key]["tile_candidates_ms"] = {k: round(v, 2) for k, v in tune["candidate_ms"].items()}
        log("sigma set %s: %.1f ms" % (key, dt * 1e3))
    engine.set_mixed_form(sel)
    if only:
        return res
    a = np.abs(outs["fp64"]).reshape(-1, 10)

    def err(k):
        d = np.abs(outs[k] - outs["fp64"]).reshape(-1, 10)
        return float((d / a.max(axis=0)).max())
    res["mixed_vs_fp64"] = round(res["mixed"]["value"] / res["fp64"]["value"], 3)
    res["max_normwise_err"] = err("mixed")
    if "bf16x3" in res:
        res["bf16x3"]["vs_fp64"] = round(res["bf16x3"]["value"] / res["fp64"]["value"], 3)
        res["bf16x3"]["max_normwise_err"] = err("bf16x3")
    res["unit"] = "correlators/s"
    res["config"] = dict(workload="sigma one_fish (4) + sigma-pi triangle (2) + rho-pi (4) per (tf, t0, p)",
                         nv=nv, lt=nt, sources=nsrc, momentum_sets=nmom, correlators_per_pass=ncorr)
    return res


def reader_measurement(nv=64, nt=48):
    """The perambulator file of readPropH (lib/contract.cpp:91-160) at Nv = 64, Lt = 48 (2.4 GB),
    written to a temporary file whose complex values are their own file index (so a sample of
    the device layout checks the permutation), then read page-cached: by lc_read_prop_device
    (slabs through pinned buffers, permuted on the GPU) and by the host restatement of readPropH
    (lc_read_prop). Runs last: the host reader's destroyContract finalises the library."""
    import ctypes
    import shutil
    import tempfile
    import numpy as np
    from contractions_code_scattering_amd import _capi, check
    from contractions_code_scattering_amd.devmem import DeviceArray, synchronize
    L = _capi.lib()
    d = tempfile.mkdtemp(dir="/tmp")
    path = os.path.join(d, "prop.bin")
    n = 16 * nv * nv * nt * nt
    try:
        with open(path, "wb") as f:
            f.write(b"synthetic".ljust(100, b"\0"))
            f.write(np.array([nt, nv], "<u4").tobytes())
            for s0 in range(0, n, 1 << 24):
                v = np.zeros(min(n, s0 + (1 << 24)) - s0, "<c16")
                v.real = np.arange(s0, s0 + v.size)
                f.write(v.tobytes())
        log("reader: wrote %.2f GB" % (n * 16 / 1e9))
        out = DeviceArray((nt, nt, 4, 4, nv, nv), np.complex128)
        res = {}
        for k in range(2):  # the first call pages the file in
            t = time.perf_counter()
            check(L.lc_read_prop_device(path.encode(), nt, nv, 1, out.ptr))
            synchronize()
            dt = time.perf_counter() - t
        res["device"] = dict(ms=round(dt * 1e3, 1), gbs=round(n * 16 / dt / 1e9, 2))
        tau = out.to_numpy()
        del out
        g = np.random.default_rng(0)
        ok = True
        for _ in range(64):
            t1, t2, v1, v2 = g.integers(0, nt), g.integers(0, nt), g.integers(0, nv), g.integers(0, nv)
            s1, s2 = g.integers(0, 4), g.integers(0, 4)
            idx = ((((s1 * nv + v1) * nt + t1) * 4 + s2) * nv + v2) * nt + t2
            ok &= tau[t1, t2, s1, s2, v1, v2] == idx
        res["permutation_sample_ok"] = bool(ok)
        del tau
        check(L.lc_contract_init(nv, nt, 0))
        try:
            ptr = ctypes.c_void_p()
            t = time.perf_counter()
            check(L.lc_read_prop(path.encode(), 1, ctypes.byref(ptr)))
            dt = time.perf_counter() - t
            L.lc_free_host(ptr)
        finally:
            check(L.lc_contract_destroy())
        res["host_readPropH"] = dict(ms=round(dt * 1e3, 1), gbs=round(n * 16 / dt / 1e9, 2))
        log("reader: device %.1f ms, host %.1f ms" % (res["device"]["ms"], res["host_readPropH"]["ms"]))
        res["config"] = dict(nv=nv, lt=nt, file_bytes=n * 16 + 108, page_cached=True)
        return res
    finally:
        shutil.rmtree(d, ignore_errors=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-rows", type=int, default=2, help="sink rows for the CPU baseline (0: skip)")
    ap.add_argument("--nmom", type=int, default=19)
    ap.add_argument("--seed", type=int, default=3, help="seed of the synthetic inputs")
    ap.add_argument("--mode", choices=["both", "fp64", "mixed"], default="both",
                    help="profiling aid: run only one form (the JSON line needs both)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="profiling aid: one device-to-device copy of this size (counter calibration)")
    ap.add_argument("--d512", type=int, default=0,
                    help="N > 0: also time rho + pi-pi at Nv=128 (D=512), Lt=64, all 64 sources, N momenta, "
                         "then N moving-frame pi-pi momentum sets and N sigma-set momentum sets on the same "
                         "perambulator (reported under d512)")
    ap.add_argument("--reader", type=int, default=0,
                    help="1: also time the perambulator file reader (GPU and host) at Nv=64, Lt=48 from a "
                         "page-cached temporary file (reported under reader; runs last)")
    ap.add_argument("--elementals", type=int, default=0,
                    help="1: also time the GPU construction of the momentum elementals at L=24, Nv=64, Lt=48, "
                         "19 momenta (reported under elementals)")
    ap.add_argument("--sigma-set", type=int, default=1,
                    help="1: also time the batched sigma / sigma-pi / rho-pi set at Nv=64, Lt=64, 16 sources, "
                         "27 momentum sets (reported under sigma_set; with --mode fp64 or mixed that form alone)")
    ap.add_argument("--bf16x3", type=int, default=0,
                    help="1: the mixed leg runs the BF16x3 form (FP32 storage, split BF16 products; it does not meet "
                         "the 1e-6 bar) instead of the exact int8-digit form")
    ap.add_argument("--autotune", type=int, default=1,
                    help="1: the mixed form tunes its tile shape on an untimed first call per workload; 0: the "
                         "built-in rule (4 tiles per workgroup at Nv=64, the tuned choice there), no extra calls "
                         "(for profiles: every traced launch is then a measured one)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path", file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))

    import numpy as np
    from contractions_code_scattering_amd import engine, rows
    from contractions_code_scattering_amd.devmem import DeviceArray, synchronize, to_device

    nv, nt, nmom = 64, 48, args.nmom
    tau_h, phi_h = workload(nv, nt, nmom, seed=args.seed)
    rows.init(0)
    # the mixed form: exact int8 digits (default) or BF16x3; the BF16x3 form picks its tile shape on
    # its first (untimed) call per shape
    engine.set_mixed_form(engine.MIXED_BF16X3 if args.bf16x3 else engine.MIXED_EXACT)
    engine.autotune(bool(args.autotune))
    tau = to_device(tau_h)
    phi = to_device(phi_h)
    nsrc = nt
    ncorr = nmom * nsrc * nt * 8
    log("inputs resident: tau %.2f GB, %d momenta" % (tau.nbytes / 1e9, nmom))

    def run(mixed):
        out = (DeviceArray((nmom, nsrc, nt, 10), np.complex128),
               DeviceArray((nmom, nsrc, nt, 8), np.complex128),
               DeviceArray((nmom, 8, nt), np.complex128))
        tune = None
        if mixed and args.autotune and args.bf16x3:
            # untimed: the autotuner times its tile-shape candidates on this call
            engine.rho_pipi(tau, phi, out=out, mixed=mixed)
            tune = engine.autotune_report()
            log("mixed tile shape: %d tiles per workgroup (%s; candidates %s ms)"
                % (tune["tiles_per_wg"], tune["how"], {k: round(v, 2) for k, v in tune["candidate_ms"].items()}))
        for i in range(args.warmup):
            engine.rho_pipi(tau, phi, out=out, mixed=mixed)
        synchronize()
        engine.profile(True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            engine.rho_pipi(tau, phi, out=out, mixed=mixed)
        synchronize()
        elapsed = time.perf_counter() - t0
        prof = engine.profile_read()
        engine.profile(False)
        log("%s: %d steps in %.3f s" % ("mixed" if mixed else "fp64", args.steps, elapsed))
        if mixed and tune is None and args.bf16x3:
            tune = engine.autotune_report()
        return dict(elapsed=elapsed, prof=prof, corr=out[1].to_numpy(), avg=out[2].to_numpy(),
                    tune=tune if mixed else None)

    if args.calib_copy_mib > 0:
        from contractions_code_scattering_amd.devmem import copy
        n = args.calib_copy_mib << 20
        src, dst = DeviceArray((n,), np.uint8), DeviceArray((n,), np.uint8)
        copy(dst, src)
        log("calibration copy of %d MiB done" % args.calib_copy_mib)
    if args.mode != "both":
        # profiling aid: one form, then only the optional legs asked for (results on stderr)
        r = run(args.mode == "mixed")
        log("single-mode run: %.1f correlators/s" % (ncorr * args.steps / r["elapsed"]))
        extra = {}
        if args.elementals:
            extra["elementals"] = elementals_measurement()
        if args.reader:
            extra["reader"] = reader_measurement()
        if args.sigma_set:
            extra["sigma_set"] = sigma_set_measurement(bf16x3=args.bf16x3, only=args.mode)
        if extra:
            log(json.dumps(extra))
        return
    r64 = run(False)
    rmx = run(True)

    def stats(r, peak):
        p = r["prof"]
        ms = r["elapsed"] * 1e3 / args.steps
        ach = p["gemm_flops"] / (p["gemm_ms"] * 1e-3) / 1e12 if p["gemm_ms"] > 0 else 0.0
        return dict(value=ncorr * args.steps / r["elapsed"], ms_per_step=ms, achieved=ach,
                    frac=ach / peak, avg_launch_ms=p["gemm_ms"] / max(p["gemm_launches"], 1),
                    flops_per_launch=p["gemm_flops"] / max(p["gemm_launches"], 1),
                    effective_tflops=p["gemm_flops"] / args.steps / (ms * 1e-3) / 1e12)

    # the roofline of the mixed form's GEMM: the exact form on the int8 matrix cores (10 int8 products
    # per algorithmic multiply-add), or the BF16x3 form on the BF16 ones (6 piece products)
    if args.bf16x3:
        mx_unit, mx_ops = "bf16", 6.0
        kname, kdesc = "lc_gemm_mx_kernel", "BF16x3 pieces on v_mfma_f32_16x16x32_bf16, FP32 accumulation"
    else:
        mx_unit, mx_ops = "int8", EXACT_OPS_PER_FLOP
        kname, kdesc = "lc_gemm_oz_kernel<8, true, *>", "int8 digits on v_mfma_i32_32x32x32_i8, exact int32 accumulation"
    peak_mx = PEAK_TFLOPS[mx_unit]
    s64 = stats(r64, PEAK_TFLOPS["f64"])
    smx = stats(rmx, peak_mx)
    errs = engine.rel_errors(rmx["avg"], r64["avg"])
    errs_t = engine.errors_per_t(rmx["avg"], r64["avg"], corr_ref=r64["corr"])
    max_rel = max(e["max_rel"] for e in errs.values())
    max_norm = max(e["max_normwise"] for e in errs.values())
    max_sig = max(e["max_rel_significant"] for e in errs.values())
    roof = dict(bound="mfma", achieved=round(smx["achieved"], 3), peak=round(peak_mx, 1), unit="TFLOP/s",
                frac=round(smx["frac"], 4), traffic=traffic_from_profiles(kname, nmom),
                traffic_source="profiles/traffic_rho_pipi_D256_Lt48_%dmom.json (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                               "scaled by a 1 GiB copy; not measured in this run)" % nmom,
                kernel=kname + " (" + kdesc + ")",
                peak_basis="published MI355X %s dense peak, %.1f %s; achieved counts algorithmic flops (8 per "
                           "complex multiply-add), not the %d %s products each costs"
                           % (mx_unit, peak_mx / 1000, "POP/s" if mx_unit == "int8" else "PFLOP/s", mx_ops, mx_unit),
                flops_per_launch=smx["flops_per_launch"], avg_launch_ms=round(smx["avg_launch_ms"], 4),
                rate_over_published_peaks={k: round(smx["achieved"] / v, 4) for k, v in PEAK_TFLOPS.items()})
    roof[mx_unit + "_issue_utilization"] = round(smx["achieved"] * mx_ops / peak_mx, 4)
    line = dict(metric=METRIC, value=round(smx["value"], 1), unit="correlators/s", n_gpus=1,
                steps=args.steps, warmup=args.warmup, ms_per_step=round(smx["ms_per_step"], 3),
                higher_is_better=True, scaling="weak", vs_baseline=None,
                dtype="bf16x3/f32/f64" if args.bf16x3 else "i8-digits/f64",
                data="synthetic",
                config=dict(workload="rho(3 pol)+pipi(5 diagrams) Nv=64 D=256 Lt=48 all 48 sources "
                                     "%d momenta" % nmom,
                            nv=nv, D=4 * nv, lt=nt, sources=nsrc, momenta=nmom, seed=args.seed,
                            mode=("mixed, BF16x3 form: FP32 storage split into three BF16 pieces, 6 piece "
                                  "products on BF16 MFMA with FP32 accumulation, FP64 traces and sums") if args.bf16x3
                            else ("mixed, exact form: 4 balanced base-255 int8 digits per value under per-line "
                                  "exponents (scale 63/32 2^(30-E)), 10 digit pairs on int8 MFMA with exact int32 "
                                  "accumulation, FP64 epilogue, closing traces and sums"),
                            correlators_per_step=ncorr),
                roofline=roof,
                mixed_vs_fp64=round(smx["value"] / s64["value"], 3),
                max_rel_err=max_rel, max_normwise_err=max_norm, max_rel_err_significant=max_sig,
                accuracy_gate=dict(bar=1e-6, against="plain FP64 engine (itself within 1e-12 of the oracle)",
                                   normwise_pass=bool(max_norm <= 1e-6),
                                   significant_elementwise_pass=bool(max_sig <= 1e-6),
                                   frac_significant_within=errs_t["frac_significant_rel_le_1e6"],
                                   frac_all_within=errs_t["frac_rel_le_1e6"]),
                errors_per_diagram=errs, errors_per_t=errs_t,
                plain_fp64=dict(value=round(s64["value"], 1), ms_per_step=round(s64["ms_per_step"], 3),
                                roofline=dict(bound="mfma", achieved=round(s64["achieved"], 3),
                                              peak=PEAK_TFLOPS["f64"], unit="TFLOP/s",
                                              frac=round(s64["frac"], 4),
                                              traffic=traffic_from_profiles("lc_gemm_kernel<double, 64>", nmom),
                                              kernel="lc_gemm_kernel<double,64>",
                                              avg_launch_ms=round(s64["avg_launch_ms"], 4))),
                effective_tflops=round(smx["effective_tflops"], 3))
    if args.sigma_set:
        line["sigma_set"] = sigma_set_measurement(bf16x3=args.bf16x3)
    if args.d512 > 0:
        line["d512"] = d512_measurement(args.d512)
    if args.elementals:
        line["elementals"] = elementals_measurement()
    if rank == 0 and args.cpu_rows > 0:
        rec, ref = cpu_baseline(tau_h, phi_h, args.cpu_rows)
        line["cpu_baseline"] = rec
        got = r64["corr"][0][:, :args.cpu_rows, :].transpose(1, 0, 2)  # [tf][t0][8]
        line["fp64_vs_oracle_max_rel"] = float(np.abs(got - ref).max() / np.abs(ref).max())
    if args.reader:
        line["reader"] = reader_measurement()
    if rank == 0:
        print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
