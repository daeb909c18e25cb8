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
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16": 2500.0}
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
    """Oracle on `rows` sink rows of momentum 0: rho op pair (0,0) + the I=0 pi-pi set."""
    import numpy as np
    import oracle as O
    from contractions_code_scattering_amd import inputs as I
    nt = tau.shape[0]
    P = phi[0]
    Pd = I.dagger(P)
    mom5 = np.ascontiguousarray(np.stack([P] * 5, axis=1))
    diag = O.diag_of(tau)
    t = time.perf_counter()
    for tf in range(rows):
        O.rho_row_range(tau[tf], mom5, tf, 0, 48)
        O.pipi_row(0, tau[tf], diag, P, Pd, Pd, P, tf)
    dt = time.perf_counter() - t
    ncorr = rows * nt * 8
    return dict(value=ncorr / dt, unit="correlators/s", cores=O.num_threads(), kind="port",
                sample="%d sink rows x %d sources x 8 correlators of momentum 0 (rho op pair (0,0) "
                       "+ pi-pi I=0 set) at Nv=64, Lt=48, FP64, %.1f s" % (rows, nt, dt))


def traffic_from_profiles():
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 PMC summary."""
    path = os.path.join(ROOT, "profiles", "gemm_f64_traffic.json")
    if os.path.exists(path):
        try:
            with open(path) as f:
                return json.load(f).get("bytes_per_launch")
        except (OSError, ValueError):
            return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-rows", type=int, default=2, help="sink rows for the CPU baseline (0: skip)")
    ap.add_argument("--nmom", type=int, default=19)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the sharding design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))

    import numpy as np
    from contractions_code_scattering_amd import engine, rows
    from contractions_code_scattering_amd.devmem import DeviceArray, synchronize, to_device

    nv, nt, nmom = 64, 48, args.nmom
    tau_h, phi_h = workload(nv, nt, nmom)
    rows.init(0)
    tau = to_device(tau_h)
    phi = to_device(phi_h)
    nsrc = nt
    out = (DeviceArray((nmom, nsrc, nt, 10), np.complex128),
           DeviceArray((nmom, nsrc, nt, 8), np.complex128),
           DeviceArray((nmom, 8, nt), np.complex128))
    log("inputs resident: tau %.2f GB, %d momenta" % (tau.nbytes / 1e9, nmom))
    for i in range(args.warmup):
        engine.rho_pipi(tau, phi, out=out)
        log("warmup step %d done" % i)
    synchronize()
    engine.profile(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        engine.rho_pipi(tau, phi, out=out)
    synchronize()
    elapsed = time.perf_counter() - t0
    log("timed %d steps: %.3f s" % (args.steps, elapsed))
    prof = engine.profile_read()
    engine.profile(False)

    ncorr = nmom * nsrc * nt * 8
    ms_per_step = elapsed * 1e3 / args.steps
    value = ncorr * args.steps / elapsed
    gemm_avg_ms = prof["gemm_ms"] / max(prof["gemm_launches"], 1)
    flops_per_launch = prof["gemm_flops"] / max(prof["gemm_launches"], 1)
    achieved = prof["gemm_flops"] / (prof["gemm_ms"] * 1e-3) / 1e12 if prof["gemm_ms"] > 0 else 0.0
    roof = dict(bound="mfma", achieved=round(achieved, 3), peak=PEAK_TFLOPS["f64"], unit="TFLOP/s",
                frac=round(achieved / PEAK_TFLOPS["f64"], 4), traffic=traffic_from_profiles(),
                kernel="lc_gemm_kernel<double,64>", flops_per_launch=flops_per_launch,
                avg_launch_ms=round(gemm_avg_ms, 4))
    line = dict(metric=METRIC, value=round(value, 1), unit="correlators/s", n_gpus=1,
                steps=args.steps, warmup=args.warmup, ms_per_step=round(ms_per_step, 3),
                higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f64",
                data="synthetic",
                config=dict(workload="rho(3 pol)+pipi(5 diagrams) Nv=64 D=256 Lt=48 all 48 sources "
                                     "%d momenta, plain FP64" % nmom,
                            nv=nv, D=4 * nv, lt=nt, sources=nsrc, momenta=nmom, mode="plain_fp64",
                            correlators_per_step=ncorr),
                roofline=roof,
                effective_tflops=round(prof["gemm_flops"] / args.steps / (ms_per_step * 1e-3) / 1e12, 3))
    if rank == 0 and args.cpu_rows > 0:
        line["cpu_baseline"] = cpu_baseline(tau_h, phi_h, args.cpu_rows)
    if rank == 0:
        print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
