"""The BF16x3 mixed GEMM's tile shape (output tiles per workgroup), the in-project autotuner, and
momentum chunking of the mixed engines.

The tile shape only regroups the same products, so every setting must give the results of the
built-in rule, to FP64 rounding of the regrouped trace partial sums (1e-13 normwise; the mixed
form itself is within 2e-6 of the FP64 engine, tested elsewhere). Shapes cover the fused-sandwich
chain (Nv = 64), the four-level chain with one tile per spin block (Nv = 40) and with four
(Nv = 128), and the sigma set.
"""
import numpy as np
import pytest

from contractions_code_scattering_amd import inputs as I
from conftest import relerr

pytestmark = pytest.mark.gpu


@pytest.fixture
def tiles(lc):
    """The engine with the BF16x3 mixed form selected (the form whose tile shape these tests set;
    the exact form has a fixed tile) and the default form restored afterwards."""
    from contractions_code_scattering_amd import engine
    engine.set_mixed_form(engine.MIXED_BF16X3)
    engine.set_tile_params(0)
    engine.autotune(False, forget=True)
    yield engine
    engine.set_tile_params(0)
    engine.autotune(False, forget=True)
    engine.set_mixed_form(engine.MIXED_EXACT)


def _rho_pipi_inputs(seed, nt, nv, np_):
    g = I.rng(seed)
    tau = I.perambulator(g, nt, nv, 0.3)
    phi = np.stack([I.random_elemental(g, nt, nv) for _ in range(np_)])
    return tau, phi


def _run(engine, t, f, **kw):
    raw, corr, avg = engine.rho_pipi(t, f, mixed=True, **kw)
    return raw.to_numpy(), corr.to_numpy(), avg.to_numpy()


@pytest.mark.parametrize("nt,nv,np_,src0,step", [(3, 64, 2, 0, 1), (4, 40, 2, 1, 2), (2, 128, 1, 0, 1)])
def test_tile_shapes_agree(tiles, nt, nv, np_, src0, step):
    from contractions_code_scattering_amd.devmem import to_device
    tau, phi = _rho_pipi_inputs(21, nt, nv, np_)
    t, f = to_device(tau), to_device(phi)
    ref = _run(tiles, t, f, src0=src0, src_step=step)
    assert tiles.autotune_report()["how"] == "fixed"
    for b in (1, 2, 4):
        tiles.set_tile_params(b)
        got = _run(tiles, t, f, src0=src0, src_step=step)
        rep = tiles.autotune_report()
        assert rep["tiles_per_wg"] == b and rep["how"] == "fixed"
        for x, y in zip(got, ref):
            assert relerr(x, y) < 1e-13, (b, relerr(x, y))


def test_autotune_tunes_once_and_keeps_results(tiles):
    from contractions_code_scattering_amd.devmem import to_device
    nt, nv, np_ = 3, 64, 2
    tau, phi = _rho_pipi_inputs(22, nt, nv, np_)
    t, f = to_device(tau), to_device(phi)
    ref = _run(tiles, t, f)
    tiles.autotune(True)
    got = _run(tiles, t, f)
    rep = tiles.autotune_report()
    print("autotune:", rep)
    assert rep["how"] == "tuned" and rep["tiles_per_wg"] in (1, 2, 4)
    assert all(ms > 0 for ms in rep["candidate_ms"].values())
    assert rep["candidate_ms"][rep["tiles_per_wg"]] == min(rep["candidate_ms"].values())
    for x, y in zip(got, ref):
        assert relerr(x, y) < 1e-13
    chosen = rep["tiles_per_wg"]
    got2 = _run(tiles, t, f)
    rep2 = tiles.autotune_report()
    assert rep2["how"] == "cached" and rep2["tiles_per_wg"] == chosen
    for x, y in zip(got2, ref):
        assert relerr(x, y) < 1e-13
    # another shape of the same engine is tuned on its own
    _run(tiles, t, f, src0=1, src_step=1)
    assert tiles.autotune_report()["how"] == "tuned"
    # an explicit override wins over the cache; switching tuning off falls back to the rule
    tiles.set_tile_params(2)
    _run(tiles, t, f)
    assert tiles.autotune_report()["tiles_per_wg"] == 2
    assert tiles.autotune_report()["how"] == "fixed"
    tiles.set_tile_params(0)
    tiles.autotune(False)
    _run(tiles, t, f)
    assert tiles.autotune_report()["how"] == "fixed" and tiles.autotune_report()["tiles_per_wg"] == 0


def test_fp64_engine_is_not_tuned(tiles):
    from contractions_code_scattering_amd.devmem import to_device
    tau, phi = _rho_pipi_inputs(23, 2, 16, 1)
    tiles.autotune(True)
    tiles.rho_pipi(to_device(tau), to_device(phi))
    assert tiles.autotune_report()["how"] == "fixed"


def test_sigma_set_tile_shapes_agree(tiles):
    from contractions_code_scattering_amd.devmem import to_device
    nt, nv, np_ = 3, 64, 1
    g = I.rng(24)
    tau = I.perambulator(g, nt, nv, 0.3)
    el = np.ascontiguousarray(np.stack([np.stack([I.random_elemental(g, nt, nv) for _ in range(11)], axis=1)
                                        for _ in range(np_)]))
    t, e = to_device(tau), to_device(el)
    raw0, loops0 = (x.to_numpy() for x in tiles.sigma_set(t, e, idir=1, mixed=True))
    for b in (1, 2):
        tiles.set_tile_params(b)
        raw, loops = (x.to_numpy() for x in tiles.sigma_set(t, e, idir=1, mixed=True))
        assert relerr(raw, raw0) < 1e-13 and relerr(loops, loops0) < 1e-13
    tiles.set_tile_params(0)
    tiles.autotune(True)
    raw, loops = (x.to_numpy() for x in tiles.sigma_set(t, e, idir=1, mixed=True))
    assert tiles.autotune_report()["how"] == "tuned"
    assert relerr(raw, raw0) < 1e-13 and relerr(loops, loops0) < 1e-13


def _one_momentum_budget(run):
    """The scratch budget of one momentum per chunk, read from the engine's own refusal."""
    import re
    from contractions_code_scattering_amd._capi import LcError
    with pytest.raises(LcError) as e:
        run(1)
    m = re.search(r"below one momentum(?: set)? \((\d+) B\)", str(e.value))
    assert m, str(e.value)
    return int(m.group(1))


def test_mixed_momentum_chunks_are_invisible(tiles):
    """One momentum per chunk (every view offset by the chunk's first momentum) gives the results
    of one chunk for all momenta, bit for bit, in the fused-sandwich rho + pi-pi chain and the
    sigma set."""
    from contractions_code_scattering_amd.devmem import to_device
    nt, nv, np_ = 3, 64, 3
    tau, phi = _rho_pipi_inputs(25, nt, nv, np_)
    t, f = to_device(tau), to_device(phi)
    whole = _run(tiles, t, f, src0=1, src_step=1)
    budget = _one_momentum_budget(lambda b: tiles.rho_pipi(t, f, mixed=True, src0=1, src_step=1,
                                                           max_scratch_bytes=b))
    chunked = _run(tiles, t, f, src0=1, src_step=1, max_scratch_bytes=budget)
    for x, y in zip(chunked, whole):
        assert np.array_equal(x, y)
    g = I.rng(26)
    el = np.ascontiguousarray(np.stack([np.stack([I.random_elemental(g, nt, nv) for _ in range(11)], axis=1)
                                        for _ in range(np_)]))
    e = to_device(el)
    raw0, loops0 = (x.to_numpy() for x in tiles.sigma_set(t, e, idir=0, mixed=True))
    budget = _one_momentum_budget(lambda b: tiles.sigma_set(t, e, idir=0, mixed=True, max_scratch_bytes=b))
    raw1, loops1 = (x.to_numpy() for x in tiles.sigma_set(t, e, idir=0, mixed=True, max_scratch_bytes=budget))
    assert np.array_equal(raw0, raw1) and np.array_equal(loops0, loops1)
