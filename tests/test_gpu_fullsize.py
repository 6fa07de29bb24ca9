"""Parity at the full sizes the throughput figures are measured on, from seeded random
initial fields (every cell, the walls and the CPML slabs live from the first step):

* the 1024^3 Yee workload (200-sphere dielectric, CPML 12, Ez point source
  and probe): the in-place engine equals the float32 C oracle after 6 steps; the fused
  engine (CE as uint8 material indices, field sets in one staggered allocation) equals
  the in-place engine (float CE) in every field, psi value and probe sample after 6, 11
  and 61 steps (58 with step pairs), with single steps and with step pairs;
* the 512^3 4th-order symplectic workload (100-sphere dielectric, CPML 10,
  S = 0.9 S_max): equal to the float32 C oracle after 6 steps;
* the 64-cube uneven batch (CPML 8): every instance of the batched run (CE as material
  indices) equals the same instance run alone (float CE) after 3 and then 33 steps.

Values are compared with -0 == +0 (the fused kernel's masked lanes may store +0 where
the in-place arithmetic gives value*0 = -0).  This is the only place the shapes that
the small parity tests cannot reach are checked: arrays of 1 GiB and more, four 1 KiB
z segments per row, the full block grids of the sweep and the shell passes."""
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
    gc.collect()
    torch.cuda.empty_cache()
    yield
    gc.collect()
    torch.cuda.empty_cache()


def _need(gb):
    free, _ = torch.cuda.mem_get_info()
    if free < gb * 1e9:
        pytest.skip(f"needs about {gb} GB of free device memory, {free / 1e9:.0f} GB free")


def _equal_host(a, b):
    a = np.ascontiguousarray(a) + np.float32(0)
    b = np.ascontiguousarray(b) + np.float32(0)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


def _yee(engine, ce_indices):
    import bench

    class Args:
        n, npml, spheres = 1024, 12, 200
    a = Args()
    a.engine = engine
    a.no_ce_lut = not ce_indices
    return bench.build_sim(a)


def _randomize(tensors, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    for t in tensors:
        t.normal_(generator=g)


def test_yee_1024cube_spheres_engines_and_oracle():
    _need(160)
    ref = _yee("kemp", False)
    fus = _yee("fused", True)
    assert ref.fields.ce_mat is None and fus.fields.ce_mat is not None
    assert ref.fields.pad == 0
    _randomize(ref.fields.ehs, 11)
    for a, b in zip(ref.fields.ehs, fus.fields.ehs):
        b.copy_(a)
    torch.cuda.synchronize()
    # float32 C oracle from the same state
    f = ref.fields
    o = O.OracleSim(f.nx, f.ny, f.nz, f.pad, ref.cpml.npml, np.float32, ces=[c.cpu().numpy() for c in f.ces])
    for k in range(6):
        o.ehs[k][...] = f.ehs[k].cpu().numpy()
    s = ref.sources[0]
    o.add_source(s.str_f, s.pt0, s.pt1, s.tfunc)
    comp, pt = ref.probe_points[0]
    o.add_probe(NAMES[comp], pt)

    a = ref.run(6)
    b = fus.run(6)
    o.run(6)
    assert _equal_host(a, np.array(o.series, np.float32))
    assert _equal_host(a, b)
    for k, name in enumerate(NAMES):
        assert _equal_host(ref.field(name).cpu().numpy(), o.ehs[k]), ("oracle", name)
        assert torch.equal(ref.field(name), fus.field(name)), ("fused", name)
    del o
    gc.collect()

    a = ref.run(5)
    b = fus.run(5)
    assert _equal_host(a, b)
    for name in NAMES:
        assert torch.equal(ref.field(name), fus.field(name)), ("fused, 11 steps", name)
    # a longer run on both engines (61 steps in all)
    a = ref.run(50)
    b = fus.run(50)
    assert _equal_host(a, b)
    for name in NAMES:
        assert torch.equal(ref.field(name), fus.field(name)), ("fused, 61 steps", name)
    for key in ref.cpml.psiE:
        for c in range(2):
            assert torch.equal(ref.cpml.psiE[key][c], fus.cpml.psiE[key][c]), ("psiE", key, c)
            assert torch.equal(ref.cpml.psiH[key][c], fus.cpml.psiH[key][c]), ("psiH", key, c)
    arrays = fus.fields.ehs + fus._fused.alt
    base = arrays[0].untyped_storage().data_ptr()
    assert all(t.untyped_storage().data_ptr() == base for t in arrays)


def test_yee_1024cube_step_pairs_equal_inplace():
    """Step pairs on the 1024^3 workload (two-step sweep over the box 14 .. 1010 in x and
    y, 16 .. 1008 in z with the field arrays at line phase 16, 1776 blocks of 12 waves;
    shell and fringe passes into the third buffer set): equal to the in-place engine
    after 2 steps (one pair), 5 more (two pairs and a single step) and 51 more, fields,
    psi and probe series."""
    _need(160)
    from fdtd_accelerate_amd.fused import FusedStepper
    ref = _yee("kemp", False)
    fus = _yee("fused", True)
    _randomize(ref.fields.ehs, 21)
    for a, b in zip(ref.fields.ehs, fus.fields.ehs):
        b.copy_(a)
    torch.cuda.synchronize()
    fus._fused = FusedStepper(fus, step_pairs=True)
    assert fus._fused.pairs and fus._fused.sweep_box() == (14, 1010, 14, 1010, 16, 1008)
    for steps in (2, 5, 51):
        a = ref.run(steps)
        b = fus.run(steps)
        assert _equal_host(a, b), steps
        for name in NAMES:
            assert torch.equal(ref.field(name), fus.field(name)), (steps, name)
    for key in ref.cpml.psiE:
        for c in range(2):
            assert torch.equal(ref.cpml.psiE[key][c], fus.cpml.psiE[key][c]), ("psiE", key, c)
            assert torch.equal(ref.cpml.psiH[key][c], fus.cpml.psiH[key][c]), ("psiH", key, c)


def test_s4_512cube_spheres_vs_oracle():
    _need(20)
    from fdtd_accelerate_amd import geometry
    from fdtd_accelerate_amd.s4 import S4Sim
    from fdtd_accelerate_amd.sources import gaussian_pulse
    n, npml, nsph = 512, 10, 100
    rmax = min(40.0, (n - 2 * npml) / 8)
    spheres = geometry.random_spheres(2024, nsph, (min(8.0, rmax / 2), rmax), (1.0, 12.0), (npml,) * 3, (n - npml,) * 3)
    eps = geometry.eps_map_numpy((n, n, n), spheres)
    g = S4Sim(n, n, n, npml=npml)
    g.set_eps(eps)
    o = O.OracleS4Sim(n, n, n, g.S, npml=npml, ces=[(np.float32(g.S / 24) / eps).astype(np.float32)] * 3)
    del eps
    c = n // 2
    tf = gaussian_pulse()
    for sim in (g, o):
        sim.add_source("ez", (c, c, c), (c, c, c), tf)
        sim.add_probe("ez", (c + n // 4, c, c))
    _randomize(g.ehs, 12)
    torch.cuda.synchronize()
    for k, name in enumerate(NAMES):
        o.ehs[k][...] = g.field(name).cpu().numpy()
    series = g.run(6)
    o.run(6)
    assert _equal_host(series, np.array(o.series, np.float32))
    for k, name in enumerate(NAMES):
        assert _equal_host(np.asfortranarray(g.field(name).cpu().numpy()), o.ehs[k]), name


def test_uneven_64cube_batch_equals_single_runs():
    _need(40)
    from fdtd_accelerate_amd import batch as B
    a = B.make_uneven_batch(seed=3, ce_table=False)  # single runs read the float CE arrays,
    b = B.make_uneven_batch(seed=3)                  # the batch the material indices + tables
    assert len(a) == 64 and all(s.fields.ce_mat is None for s in a)
    assert all(s.fields.ce_mat is not None for s in b)
    for n, (sa, sb) in enumerate(zip(a, b)):
        _randomize(sa.fields.ehs, 100 + n)
        for x, y in zip(sa.fields.ehs, sb.fields.ehs):
            y.copy_(x)
    torch.cuda.synchronize()
    series_a = [np.concatenate([s.run(3), s.run(30)]) for s in a]
    bt = B.BatchedYee(b)
    series_b = [np.concatenate(x) for x in zip(bt.run(3), bt.run(30))]
    torch.cuda.synchronize()
    for n, (sa, sb) in enumerate(zip(a, b)):
        assert _equal_host(series_a[n], series_b[n]), n
        for name in NAMES:
            assert torch.equal(sa.field(name), sb.field(name)), (n, name)
        for key in sa.cpml.psiE:
            for c in range(2):
                assert torch.equal(sa.cpml.psiE[key][c], sb.cpml.psiE[key][c]), (n, "psiE", key, c)
                assert torch.equal(sa.cpml.psiH[key][c], sb.cpml.psiH[key][c]), (n, "psiH", key, c)
    bt.close()
