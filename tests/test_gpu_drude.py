"""Drude dispersion on the GPU (``Drude``, ``mifdtd_drude_e``) against the float32
oracle: every field, the probe series and the auxiliary current byte-equal (-0 == +0)
after the steps -- a Drude box with per-cell plasma frequency and damping inside a
dielectric grid with CPML and a point source (inside the CPML frame, over the whole grid
and across two CPML slabs, the latter two from random fields), a whole-grid uniform medium with periodic
boundaries (the plasma oscillation), the same boxes on the fused engine (one Drude box in its sweep and shell E pass), and the
engines that reject a Drude medium."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

NAMES = ("ex", "ey", "ez", "hx", "hy", "hz")


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from fdtd_accelerate_amd import _lib
    _lib.load()


def _same(a, b):
    return a.shape == b.shape and (np.ascontiguousarray(a) + np.float32(0)).tobytes() == \
        (np.ascontiguousarray(b) + np.float32(0)).tobytes()


@pytest.mark.parametrize("ns,npml,p0,p1,random_init,engine,fp", [
    ((26, 24, 29), 4, (9, 8, 10), (17, 15, 19), False, "kemp", None),   # inside the CPML frame
    ((26, 24, 29), 4, (0, 0, 0), (25, 23, 28), True, "kemp", None),     # whole grid: walls, every slab
    ((26, 24, 29), 4, (0, 5, 20), (6, 23, 28), True, "kemp", None),     # corner across x-front, z-back
    # fused engine: the sweep box is z 32..64 of nz = 100 (CPML 4), so Drude cells sit in the
    # sweep (owned and recomputed halo rows / seed planes) and in the shell E pass
    ((26, 24, 100), 4, (9, 8, 20), (17, 15, 70), False, "fused", {"rows_per_thread": 3, "planes_per_chunk": 5}),
    ((26, 24, 100), 4, (0, 0, 0), (25, 23, 99), True, "fused", {"rows_per_thread": 2, "planes_per_chunk": 4}),
    ((26, 24, 100), 4, (0, 5, 40), (6, 23, 99), True, "fused", {"rows_per_thread": 7, "planes_per_chunk": 64}),
    ((30, 28, 300), 5, (7, 3, 30), (20, 26, 290), True, "fused", None),
    ((20, 22, 70), 0, (0, 0, 0), (19, 21, 69), True, "fused", {"rows_per_thread": 4, "planes_per_chunk": 3}),
    ((26, 24, 29), 4, (0, 5, 20), (6, 23, 28), True, "fused", None),    # no sweep box: shell only
    ((26, 24, 100), 4, (3, 2, 10), (20, 21, 90), True, "fused_table", None),  # CE as material indices
    ((30, 26, 132), 4, (6, 5, 36), (20, 19, 83), True, "fused", None),  # box rows of 48 cells from k = 36
], ids=["interior", "whole_grid", "cpml_corner", "fused_interior", "fused_whole_grid", "fused_corner_T7",
        "fused_2seg", "fused_no_cpml", "fused_shell_only", "fused_ce_table", "fused_aligned_rows"])
def test_drude_box_matches_oracle(ns, npml, p0, p1, random_init, engine, fp):
    from fdtd_accelerate_amd import Drude
    from fdtd_accelerate_amd.sim import YeeSim
    from fdtd_accelerate_amd.sources import gaussian_pulse
    steps = 60
    shape = tuple(b - a + 1 for a, b in zip(p0, p1))
    rng = np.random.default_rng(3)
    ces = [(0.5 / rng.uniform(1, 4, ns)).astype(np.float32) for _ in range(3)]
    table = engine == "fused_table"
    if table:  # few distinct values: CE held as uint8 indices into the table
        engine = "fused"
        vals = (0.5 / np.array([1.0, 2.25, 4.0])).astype(np.float32)
        ces = [vals[rng.integers(0, 3, ns)] for _ in range(3)]
    wp = [rng.uniform(0.2, 1.5, shape).astype(np.float32) for _ in range(3)]
    gamma = [rng.uniform(0.0, 0.3, shape).astype(np.float32) for _ in range(3)]
    init = [rng.standard_normal(ns).astype(np.float32) for _ in range(6)] if random_init else None
    sim = YeeSim(*ns, npml=npml, coeff_use="e", engine=engine, fused_params=fp)
    f = sim.fields
    f.set_ces(*ces)
    if init is not None:
        f.set_ehs(*init)
    d = Drude(f, p0, p1, wp, gamma)
    if table:
        assert f.compress_ce()
    tf = gaussian_pulse(10.0, 4.0, 12.0)
    src, prb = (13, 12, 14), (15, 12, 14)
    if ns[2] > 40:  # source and probe inside the fused engine's sweep box
        src, prb = (13, 12, 45), (15, 12, 46)
    sim.add_source("ez", src, src, tf)
    sim.add_probe("ez", prb)
    pad = f.pad
    o = O.OracleSim(*ns, pad, npml, np.float32,
                    ces=[np.pad(c, ((0, 0), (0, 0), (0, pad)), constant_values=np.float32(0.5)) for c in ces])
    if init is not None:
        for k in range(6):
            o.ehs[k][:, :, :ns[2]] = init[k]
    cja, cjb = zip(*[O.drude_coefficients(w, g, f.dt) for w, g in zip(wp, gamma)])
    od = O.OracleDrude(p0, p1, cja, cjb)
    o.drude.append(od)
    o.add_source("ez", src, src, tf)
    o.add_probe("ez", prb)
    # odd then even call lengths: the fused engine swaps field and current sets on parity
    series = np.concatenate([sim.run(steps // 2 + 1), sim.run(steps // 2 - 1)])
    if engine == "fused":
        assert sim._fused.drude is d
    o.run(steps)
    assert _same(series, np.array(o.series, np.float32))
    for k, name in enumerate(NAMES):
        assert _same(sim.field(name).cpu().numpy(), o.ehs[k]), name
    for c in range(3):
        assert _same(d.J[c].cpu().numpy(), od.J[c]), ("J", c)
    assert np.abs(od.J[2]).max() > 1e-3


def test_drude_uniform_periodic_plasma_oscillation():
    from fdtd_accelerate_amd import Drude, Pbc
    from fdtd_accelerate_amd.sim import YeeSim
    n, wp, steps = 9, 0.6, 80
    sim = YeeSim(n, n, n, npml=0, coeff_use="")
    f = sim.fields
    f.set_ehs(*[np.ones((n, n, n), np.float32)] * 3, *[np.zeros((n, n, n), np.float32)] * 3)
    Drude(f, (0, 0, 0), (n - 1,) * 3, wp, 0.0)
    Pbc(f, "xyz")
    sim.add_probe("ez", (4, 4, 4))
    series = sim.run(steps)[:, 0]
    o = O.OracleSim(n, n, n, f.pad, 0, np.float32, pbc="xyz")
    for k in range(3):
        o.ehs[k][:, :, :n] = 1.0
    cja, cjb = O.drude_coefficients(wp, 0.0, f.dt)
    o.drude.append(O.OracleDrude((0, 0, 0), (n - 1,) * 3, [cja] * 3, [cjb] * 3))
    o.add_probe("ez", (4, 4, 4))
    o.run(steps)
    assert _same(series, np.array(o.series, np.float32)[:, 0])
    # float32 plasma oscillation: e(n+2) = (2 - (wp dt)^2) e(n+1) - e(n) up to round-off
    c = 2.0 - (wp * f.dt) ** 2
    s = series.astype(np.float64)
    assert np.abs(s[2:] - (c * s[1:-1] - s[:-2])).max() < 1e-5


def test_drude_arguments_and_engines():
    from fdtd_accelerate_amd import Drude
    from fdtd_accelerate_amd.batch import BatchedYee
    from fdtd_accelerate_amd.sim import YeeSim
    s = YeeSim(16, 16, 16, npml=0, coeff_use="")
    with pytest.raises(ValueError):
        Drude(s.fields, (0, 0, 0), (16, 3, 3), 1.0)
    Drude(s.fields, (2, 2, 2), (5, 5, 5), 1.0, 0.1)
    with pytest.raises(ValueError, match="Drude"):
        BatchedYee([s])
    s2 = YeeSim(16, 16, 16, npml=0, coeff_use="", engine="slab")
    Drude(s2.fields, (2, 2, 2), (5, 5, 5), 1.0)
    with pytest.raises(ValueError, match="Drude"):
        s2.run(1)
    # the fused engine carries one box, not two, and no step pairs with it
    s3 = YeeSim(16, 16, 64, npml=4, coeff_use="", engine="fused")
    Drude(s3.fields, (2, 2, 2), (5, 5, 5), 1.0)
    Drude(s3.fields, (7, 7, 7), (9, 9, 9), 1.0)
    with pytest.raises(ValueError, match="Drude"):
        s3.run(1)
    s4 = YeeSim(16, 16, 64, npml=4, coeff_use="", engine="fused", fused_params={"step_pairs": True})
    Drude(s4.fields, (2, 2, 2), (5, 5, 5), 1.0)
    with pytest.raises(ValueError, match="Drude"):
        s4.run(1)
