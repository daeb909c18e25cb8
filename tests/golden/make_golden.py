"""Writes the golden fixtures: seeded inputs and the oracle's outputs for every sink-row diagram.

    python tests/golden/make_golden.py

One file per size, ``rows_nv{Nv}_nt{Nt}.npz`` (numpy, no pickled objects), holding the inputs
(perambulator tau [Nt][Nt][4][4][Nv][Nv], the elementals) and, for every sink row tf, the
outputs of the oracle's row restatement of

    rho (run_ContractRho)            [Nt][Nt][16][3]
    rhopi (run_ContractRhoPi), idir   [3][Nt][Nt][4]
    pipi I=1 and I=0                  [2][Nt][Nt][5]
    loop_unity                        [Nt]
    sigma one_fish                    [Nt][Nt][4]
    sigmapi hor_fish                  [Nt]
    sigmapi triangle                  [Nt][Nt][2]

The GPU parity tests compare the HIP rows with these arrays, so they need neither the oracle
library nor the generator on the machine that runs them; tests/test_golden.py re-runs the oracle
on the stored inputs to keep the fixtures and the oracle in step.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

SIZES = [(4, 4, 11), (8, 4, 12)]  # (nv, nt, seed)


def fixture_path(nv, nt):
    return os.path.join(HERE, "rows_nv%d_nt%d.npz" % (nv, nt))


def make_inputs(nv, nt, seed):
    from contractions_code_scattering_amd import inputs as I
    g = I.rng(seed)
    tau = I.perambulator(g, nt, nv, 0.3)
    mats = np.stack([I.random_elemental(g, nt, nv) for _ in range(4)])
    mom5 = np.ascontiguousarray(np.stack([I.hermitian_elemental(g, nt, nv) for _ in range(5)], axis=1))
    x2 = np.ascontiguousarray(np.stack([I.random_elemental(g, nt, nv) for _ in range(2)], axis=1))
    return dict(tau=tau, mats=mats, mom5=mom5, x2=x2)


def oracle_outputs(inp):
    import oracle as O
    tau, mats, mom5, x2 = inp["tau"], inp["mats"], inp["mom5"], inp["x2"]
    nt = tau.shape[0]
    diag = O.diag_of(tau)
    p1, p2, p3, p4 = mats
    out = dict(
        rho=np.stack([O.rho_row(tau[tf], mom5, tf) for tf in range(nt)]),
        rhopi=np.stack([np.stack([O.rhopi_row(tau[tf], diag, p1, p2, mom5, tf, idir) for tf in range(nt)])
                        for idir in range(3)]),
        pipi=np.stack([np.stack([O.pipi_row(iso, tau[tf], diag, p1, p2, p3, p4, tf) for tf in range(nt)])
                       for iso in (1, 0)]),
        loop_unity=O.loop_unity(diag, p1),
        sigma=np.stack([O.sigma_row(tau[tf], x2, tf) for tf in range(nt)]),
        sigmapi_fish_hor=O.sigmapi_fish_hor(diag, p1, p2),
        sigmapi_triangle=np.stack([O.sigmapi_triangle_row(tau[tf], diag, p1, p2, x2, tf) for tf in range(nt)]),
    )
    return out


def main():
    for nv, nt, seed in SIZES:
        inp = make_inputs(nv, nt, seed)
        out = oracle_outputs(inp)
        np.savez_compressed(fixture_path(nv, nt), seed=seed, **inp, **{"out_" + k: v for k, v in out.items()})
        print("wrote", fixture_path(nv, nt), os.path.getsize(fixture_path(nv, nt)), "bytes")


if __name__ == "__main__":
    main()
