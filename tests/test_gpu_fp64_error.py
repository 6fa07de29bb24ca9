"""Error of the float32 GPU runs against the float64 oracle, per workload, over each
workload's full step count, at sizes the host oracle holds (the throughput runs'
1024^3 / 512^3 grids do not fit a float64 CPU run; their media, CPML and sources are
kept, the grids reduced):

* the 1024^3 Yee workload's medium (spheres eps_r U[1,12], CPML 12, Ez point source and
  probe), 160^3, 500 steps, fused engine (CE as material indices);
* the 512^3 4th-order workload's medium (100 spheres, CPML 10, S = 0.9 S_max), 96^3,
  400 steps;
* the uneven batch's instances (their own sphere maps, CPML 8, seeded sources), the
  first three instances of the seed-3 batch, 300 steps in one batched run.

Each run is first checked against the float32 oracle (bit for bit, -0 == +0), then its
probe series and fields against the float64 oracle on the same inputs (the float32
coefficient values, float64 arithmetic): max |f32 - f64| / max |f64|, printed and
bounded by 1e-4.  The dipole workload (128^3, 1000 steps) is covered in
test_gpu_yee.py."""
import gc

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

NAMES = ("ex", "ey", "ez", "hx", "hy", "hz")


@pytest.fixture(autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from fdtd_accelerate_amd import _lib
    _lib.load()
    yield
    gc.collect()
    torch.cuda.empty_cache()


def _canon(a):
    return (np.ascontiguousarray(a, np.float32) + np.float32(0)).tobytes()


def _rel(a32, a64):
    a64 = np.asarray(a64, np.float64)
    return float(np.abs(np.asarray(a32, np.float64) - a64).max() / np.abs(a64).max())


def _report(name, series32, series64, fields32, fields64):
    ps = _rel(series32, series64)
    fe = max(_rel(fields32[k], fields64[k]) for k in range(3))
    fh = max(_rel(fields32[k], fields64[k]) for k in range(3, 6))
    print(f"{name}: probe max-rel vs fp64 {ps:.3e}; E max err / max|E| {fe:.3e}; H max err / max|H| {fh:.3e}",
          flush=True)
    assert ps < 1e-4 and fe < 1e-4 and fh < 1e-4
    return ps, fe, fh


def _yee_oracle(sim, dtype, steps):
    """The oracle of a YeeSim's medium, CPML, sources (from tstep 1) and probes, run for
    ``steps`` from zero fields."""
    f = sim.fields
    o = O.OracleSim(f.nx, f.ny, f.nz, f.pad, sim.cpml.npml, dtype,
                    ces=[c.cpu().numpy().astype(dtype) for c in f.ces])
    for s in sim.sources:
        o.add_source(s.str_f, s.pt0, s.pt1, s.tfunc)
    for comp, pt in sim.probe_points:
        o.add_probe(NAMES[comp], pt)
    o.run(steps)
    return o


def test_yee_sphere_medium_160cube_500_steps():
    from fdtd_accelerate_amd import geometry
    from fdtd_accelerate_amd.sim import YeeSim
    from fdtd_accelerate_amd.sources import gaussian_pulse
    n, npml, steps = 160, 12, 500
    sim = YeeSim(n, n, n, npml=npml, coeff_use="e", engine="fused")
    spheres = geometry.random_spheres(2024, 200, (2.0, 8.0), (1.0, 12.0), (npml,) * 3, (n - npml,) * 3)
    geometry.paint_ce(sim.fields.ces, spheres, (n, n, n))
    c = n // 2
    sim.add_source("ez", (c, c, c), (c, c, c), gaussian_pulse())
    sim.add_probe("ez", (c + n // 4, c, c))
    assert sim.fields.compress_ce()
    series = sim.run(steps)[:, 0]
    got = [sim.field(nm).cpu().numpy() for nm in NAMES]
    o32 = _yee_oracle(sim, np.float32, steps)  # the oracle's sources start at tstep 1 too
    assert _canon(series) == _canon(np.array(o32.series, np.float32)[:, 0])
    for k, nm in enumerate(NAMES):
        assert _canon(got[k]) == _canon(o32.ehs[k]), nm
    del o32
    o64 = _yee_oracle(sim, np.float64, steps)
    _report("sphere Yee 160^3 CPML 12, 500 steps", series, np.array(o64.series)[:, 0], got, o64.ehs)
    assert np.abs(series).max() > 0


def test_s4_sphere_medium_96cube_400_steps():
    from fdtd_accelerate_amd import geometry
    from fdtd_accelerate_amd.s4 import S4Sim
    from fdtd_accelerate_amd.sources import gaussian_pulse
    n, npml, nsph, steps = 96, 10, 100, 400
    spheres = geometry.random_spheres(2024, nsph, (2.0, 5.0), (1.0, 12.0), (npml,) * 3, (n - npml,) * 3)
    eps = geometry.eps_map_numpy((n, n, n), spheres)
    g = S4Sim(n, n, n, npml=npml)
    g.set_eps(eps)
    ce32 = (np.float32(g.S / 24) / eps).astype(np.float32)
    c = n // 2
    tf = gaussian_pulse()
    g.add_source("ez", (c, c, c), (c, c, c), tf)
    g.add_probe("ez", (c + n // 4, c, c))
    series = g.run(steps)[:, 0]
    got = [np.asfortranarray(g.field(nm).cpu().numpy()) for nm in NAMES]
    outs = {}
    for dtype in (np.float32, np.float64):
        o = O.OracleS4Sim(n, n, n, g.S, npml=npml, dtype=dtype, ces=[ce32.astype(dtype)] * 3)
        o.add_source("ez", (c, c, c), (c, c, c), tf)
        o.add_probe("ez", (c + n // 4, c, c))
        o.run(steps)
        outs[dtype] = o
    o32 = outs[np.float32]
    assert _canon(series) == _canon(np.array(o32.series, np.float32)[:, 0])
    for k, nm in enumerate(NAMES):
        assert _canon(got[k]) == _canon(o32.ehs[k]), nm
    o64 = outs[np.float64]
    _report(f"4th-order symplectic 96^3 CPML 10, S={g.S:.4f}, 400 steps", series, np.array(o64.series)[:, 0], got,
            o64.ehs)


def test_uneven_batch_instances_300_steps():
    from fdtd_accelerate_amd import batch as B
    steps = 300
    sims = B.make_uneven_batch(seed=3, count=3)
    bt = B.BatchedYee(sims)
    out = bt.run(steps)
    bt.close()
    torch.cuda.synchronize()
    for n, sim in enumerate(sims):
        got = [sim.field(nm).cpu().numpy() for nm in NAMES]
        series = out[n][:, 0]
        o32 = _yee_oracle(sim, np.float32, steps)
        assert _canon(series) == _canon(np.array(o32.series, np.float32)[:, 0]), n
        for k, nm in enumerate(NAMES):
            assert _canon(got[k]) == _canon(o32.ehs[k]), (n, nm)
        del o32
        o64 = _yee_oracle(sim, np.float64, steps)
        f = sim.fields
        _report(f"batch instance {n} ({f.nx}^3, CPML {sim.cpml.npml}), 300 steps", series,
                np.array(o64.series)[:, 0], got, o64.ehs)
