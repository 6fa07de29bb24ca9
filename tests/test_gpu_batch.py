"""GPU parity of the batched engine: every instance of an uneven batch equals the same
simulation run alone on the in-place engine, byte for byte (fields, psi, probe
series), with and without hipGraph capture and XCD block runs; plus the oracle for one
instance."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from fdtd_accelerate_amd import _lib
    _lib.load()


def _mk(shapes, npml, seed, ce=True, materials=False):
    """Instances with random fields, CE (random values, or with ``materials`` a few
    values per instance), two sources and two probes."""
    from fdtd_accelerate_amd.sim import YeeSim
    from fdtd_accelerate_amd.sources import gaussian_pulse
    rng = np.random.default_rng(seed)
    sims = []
    for n, (nx, ny, nz) in enumerate(shapes):
        s = YeeSim(nx, ny, nz, npml=npml, coeff_use="e" if ce else "")
        f = s.fields
        f.set_ehs(*[rng.standard_normal((nx, ny, nz)).astype(np.float32) for _ in range(6)])
        if ce and materials:
            vals = (0.5 / rng.uniform(1, 12, 3 + n)).astype(np.float32)
            f.set_ces(*[vals[rng.integers(0, len(vals), (nx, ny, nz))] for _ in range(3)])
        elif ce:
            f.set_ces(*[(0.5 / rng.uniform(1, 6, (nx, ny, nz))).astype(np.float32) for _ in range(3)])
        tf = gaussian_pulse(3.0 + n, 2.0, 7.0)
        s.add_source("ez", (nx // 2, ny // 2, nz // 2), (nx // 2, ny // 2, nz // 2), tf)
        if n % 2:
            s.add_source("hx", (2, 1, 1), (2, ny - 2, nz - 2), tf, 0.5)
        s.add_probe("ez", (nx // 2 + 1, ny // 2, nz // 2))
        s.add_probe("hy", (nx - 1, 3, 2))
        sims.append(s)
    torch.cuda.synchronize()
    return sims


SHAPES = [(24, 20, 32), (30, 26, 40), (17, 23, 45), (40, 36, 60), (12, 13, 100), (26, 70, 30)]


@pytest.mark.parametrize("use_graph,remap", [(False, 0), (True, 0), (False, 4), (True, 16)])
def test_batch_matches_single_runs(use_graph, remap):
    from fdtd_accelerate_amd.batch import BatchedYee
    steps = 7
    a = _mk(SHAPES, 4, 1)
    b = _mk(SHAPES, 4, 1)
    series_a = [s.run(steps) for s in a]
    bt = BatchedYee(b, use_graph=use_graph, remap_rows=remap)
    series_b = bt.run(steps)
    torch.cuda.synchronize()
    for n, (sa, sb) in enumerate(zip(a, b)):
        for name in ("ex", "ey", "ez", "hx", "hy", "hz"):
            assert sa.field(name).cpu().numpy().tobytes() == sb.field(name).cpu().numpy().tobytes(), (n, name)
        assert series_a[n].tobytes() == series_b[n].tobytes(), n
        for key in sa.cpml.psiE:
            for c in range(2):
                assert sa.cpml.psiE[key][c].cpu().numpy().tobytes() == sb.cpml.psiE[key][c].cpu().numpy().tobytes()
                assert sa.cpml.psiH[key][c].cpu().numpy().tobytes() == sb.cpml.psiH[key][c].cpu().numpy().tobytes()
    bt.close()


@pytest.mark.parametrize("compressed", ["all", "some"])
def test_batch_ce_tables_match_single_runs(compressed):
    """CE as uint8 material indices + a table per instance (Fields.compress_ce): the batch
    reads them when every instance has them, else the float arrays of all; either way
    byte-equal to the instances run alone on the in-place engine with float CE arrays."""
    from fdtd_accelerate_amd.batch import BatchedYee
    steps = 6
    a = _mk(SHAPES, 4, 2, materials=True)
    b = _mk(SHAPES, 4, 2, materials=True)
    for n, s in enumerate(b):
        if compressed == "all" or n % 2:
            assert s.fields.compress_ce()
    series_a = [s.run(steps) for s in a]
    bt = BatchedYee(b)
    series_b = bt.run(steps)
    torch.cuda.synchronize()
    for n, (sa, sb) in enumerate(zip(a, b)):
        for name in ("ex", "ey", "ez", "hx", "hy", "hz"):
            assert sa.field(name).cpu().numpy().tobytes() == sb.field(name).cpu().numpy().tobytes(), (n, name)
        assert series_a[n].tobytes() == series_b[n].tobytes(), n
        for key in sa.cpml.psiE:
            for c in range(2):
                assert sa.cpml.psiE[key][c].cpu().numpy().tobytes() == sb.cpml.psiE[key][c].cpu().numpy().tobytes()
    bt.close()


def test_batch_follows_buffers_moved_by_another_engine():
    """An instance of the batch run alone on the fused engine between batched runs (which
    carves its field sets anew and swaps its psi buffers): the batch builds its device
    descriptors again and every instance stays byte-equal to the same simulation run
    alone in place; a Drude box added to an instance is refused at the next run."""
    from fdtd_accelerate_amd import Drude
    from fdtd_accelerate_amd.batch import BatchedYee
    a = _mk(SHAPES[:3], 4, 3)
    b = _mk(SHAPES[:3], 4, 3)
    bt = BatchedYee(b)
    bt.run(3)
    for s in a:
        s.run(3)
    b[1].engine = "fused"
    b[1].run(3)  # odd: the state ends in the fused engine's other field set
    a[1].run(3)
    b[1].engine = "kemp"
    bt.run(4)
    for s in a:
        s.run(4)
    torch.cuda.synchronize()
    for n, (sa, sb) in enumerate(zip(a, b)):
        for name in ("ex", "ey", "ez", "hx", "hy", "hz"):
            assert sa.field(name).cpu().numpy().tobytes() == sb.field(name).cpu().numpy().tobytes(), (n, name)
        for key in sa.cpml.psiE:
            for c in range(2):
                assert sa.cpml.psiE[key][c].cpu().numpy().tobytes() == sb.cpml.psiE[key][c].cpu().numpy().tobytes()
                assert sa.cpml.psiH[key][c].cpu().numpy().tobytes() == sb.cpml.psiH[key][c].cpu().numpy().tobytes()
    nx, ny, nz = SHAPES[0]
    Drude(b[0].fields, (6, 6, 6), (nx - 7, ny - 7, nz - 7), 0.5, 0.05)
    with pytest.raises(ValueError, match="Drude"):
        bt.run(1)
    bt.close()


def test_batch_continues_and_no_cpml():
    """Two run() calls (source time continues) on a CPML-free, CE-free batch."""
    from fdtd_accelerate_amd.batch import BatchedYee
    a = _mk(SHAPES[:3], 0, 5, ce=False)
    b = _mk(SHAPES[:3], 0, 5, ce=False)
    for s in a:
        s.run(9)
    bt = BatchedYee(b)
    bt.run(4)
    bt.run(5)
    torch.cuda.synchronize()
    for sa, sb in zip(a, b):
        for name in ("ex", "ey", "ez", "hx", "hy", "hz"):
            assert sa.field(name).cpu().numpy().tobytes() == sb.field(name).cpu().numpy().tobytes(), name
        assert sa.series().tobytes() == sb.series().tobytes()


def test_batch_instance_vs_oracle():
    """One instance of a batch (128^3 vacuum, CPML 10, the dipole run's pulse source) against the
    float32 oracle over 300 steps, next to two other instances."""
    from fdtd_accelerate_amd.batch import BatchedYee
    from fdtd_accelerate_amd.sim import YeeSim
    from fdtd_accelerate_amd.sources import gaussian_pulse
    sims = [YeeSim(n, n, n, npml=10, coeff_use="e") for n in (64, 128, 96)]
    for s in sims:
        c = s.fields.nx // 2
        s.add_source("ez", (c, c, c), (c, c, c), gaussian_pulse())
        s.add_probe("ez", (c + s.fields.nx // 4, c, c))
    bt = BatchedYee(sims)
    out = bt.run(300)
    o = O.OracleSim(128, 128, 128, 0, 10, np.float32)
    o.add_source("ez", (64, 64, 64), (64, 64, 64), O.pulse)
    o.add_probe("ez", (96, 64, 64))
    o.run(300)
    assert out[1][:, 0].tobytes() == np.array(o.series, np.float32)[:, 0].tobytes()
    for name in ("ex", "ey", "ez", "hx", "hy", "hz"):
        assert sims[1].field(name).cpu().numpy().tobytes() == o.field(name).tobytes(), name


def test_batch_rejects_mixed():
    from fdtd_accelerate_amd.batch import BatchedYee
    from fdtd_accelerate_amd.sim import YeeSim
    with pytest.raises(ValueError):
        BatchedYee([YeeSim(24, 24, 24, npml=4), YeeSim(24, 24, 24, npml=6)])
    with pytest.raises(ValueError):
        BatchedYee([YeeSim(24, 24, 24, coeff_use="e"), YeeSim(24, 24, 24, coeff_use="")])
