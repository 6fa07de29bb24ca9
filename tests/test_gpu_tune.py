"""GPU tests of the autotuners (sweep tiles, slab depth, batch launch options).

Each tuner times its candidates on the simulation itself, skips the candidates that do
not apply, returns the fastest and records it in memory only (save=False: the shipped
tables are not written). Tuning runs real steps, so afterwards the simulation must hold
the state of the same number of steps of the in-place engine: every field, psi and
probe value bit-identical (the fused sweep may store +0 where the in-place path stores
-0 in a masked lane, so values are compared after adding +0)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

NAMES = ("ex", "ey", "ez", "hx", "hy", "hz")


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from fdtd_accelerate_amd import _lib
    _lib.load()


@pytest.fixture
def tables():
    """Restore the in-memory tile and slab tables the tuners update."""
    from fdtd_accelerate_amd import autotune
    tiles, slabs = dict(autotune._load()), dict(autotune._load_slab())
    yield
    autotune._table, autotune._slab_table = tiles, slabs


def _sim(ns, npml, engine, seed=4):
    from fdtd_accelerate_amd.sim import YeeSim
    from fdtd_accelerate_amd.sources import gaussian_pulse
    rng = np.random.default_rng(seed)
    s = YeeSim(*ns, npml=npml, coeff_use="e", engine=engine)
    s.fields.set_ehs(*[rng.standard_normal(ns).astype(np.float32) for _ in range(6)])
    s.fields.set_ces(*[(0.5 / rng.uniform(1, 6, ns)).astype(np.float32) for _ in range(3)])
    c = tuple(n // 2 for n in ns)
    s.add_source("ez", c, c, gaussian_pulse(4.0, 3.0, 8.0))
    return s


def _canon(t):
    return (np.ascontiguousarray(t.cpu().numpy(), np.float32) + np.float32(0.0)).tobytes()


def _assert_same(sim, ref):
    torch.cuda.synchronize()
    for n in NAMES:
        assert _canon(sim.field(n)) == _canon(ref.field(n)), n
    if ref.cpml is not None:
        for key in ref.cpml.psiE:
            for c in range(2):
                assert _canon(sim.cpml.psiE[key][c]) == _canon(ref.cpml.psiE[key][c]), ("psiE", key, c)
                assert _canon(sim.cpml.psiH[key][c]) == _canon(ref.cpml.psiH[key][c]), ("psiH", key, c)


def test_tune_fused_tiles(tables):
    from fdtd_accelerate_amd import autotune
    ns, npml, steps = (44, 40, 100), 4, 2
    fits = ((2, 8), (3, 16), (11, 8))      # T = 11 runs the 1024-thread build
    cands = fits + ((9, 4096),)            # needs more than 160 KiB of LDS: skipped
    sim = _sim(ns, npml, "fused")
    best, res = autotune.tune_fused(sim, candidates=cands, steps=steps, save=False)
    assert best in fits and set(res) == set(fits) and all(v > 0 for v in res.values())
    f = sim.fields
    assert autotune.fused_tiles(f.nx, f.ny, f.nz_pitch, npml) == best
    ref = _sim(ns, npml, "kemp")
    ref.run(len(fits) * (steps + 1))
    _assert_same(sim, ref)
    # the engine's own stepper (built on the first run) takes the tuned tile and
    # continues from the tuned state
    sim.run(3)
    ref.run(3)
    assert (sim._fused.T, sim._fused.CX) == best
    _assert_same(sim, ref)
    with pytest.raises(ValueError, match="LDS"):
        autotune.tune_fused(sim, candidates=((9, 4096),), steps=1, save=False)


def test_tune_slab_depth(tables):
    from fdtd_accelerate_amd import autotune
    ns, npml, steps = (30, 26, 40), 4, 2
    sim = _sim(ns, npml, "slab")
    best, res = autotune.tune_slab(sim, candidates=(1, 2, 7, 64), steps=steps, save=False)  # 64 > nx: skipped
    assert best in (1, 2, 7) and set(res) == {1, 2, 7}
    f = sim.fields
    assert autotune.slab_planes(f.nx, f.ny, f.nz_pitch, npml, f.ce_on) == best
    ref = _sim(ns, npml, "kemp")
    ref.run(3 * (steps + 1))
    _assert_same(sim, ref)
    with pytest.raises(ValueError, match="slab depth"):
        autotune.tune_slab(sim, candidates=(64,), steps=1, save=False)


def test_tune_batch_options():
    from fdtd_accelerate_amd import batch as B
    shapes = [(24, 20, 32), (30, 26, 40), (17, 23, 45)]
    sims = [_sim(s, 4, "kemp", seed=n) for n, s in enumerate(shapes)]
    refs = [_sim(s, 4, "kemp", seed=n) for n, s in enumerate(shapes)]
    bt = B.BatchedYee(sims)
    try:
        cands = ((False, 0), (True, 0), (False, 4))
        steps = 2
        best, res = B.tune(bt, candidates=cands, steps=steps, save=False)
        assert best in cands and set(res) == set(cands) and all(v > 0 for v in res.values())
        assert (bt.use_graph, bt.remap_rows) == best
        with pytest.raises(ValueError):
            B.tune(bt, candidates=(), save=False)
    finally:
        bt.close()
    for r in refs:
        r.run(len(cands) * (steps + 2))
    for s, r in zip(sims, refs):
        _assert_same(s, r)
