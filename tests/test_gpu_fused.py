"""GPU parity of the fused single-pass step against the oracle and the in-place path.

Bar: every value bit-identical (compared as bytes after mapping -0.0 to +0.0: the
fused kernel may store +0 where the in-place path stores value*0 = -0 in a masked lane),
for every tile shape, chunk length, odd/even step counts and grid edge cases."""
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


def canon(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return (a + np.float32(0.0)).tobytes()


def same_values(a, b):
    return np.shape(a) == np.shape(b) and canon(a) == canon(b)


def _sim(ns, npml, engine, T=None, CX=None, ce=True):
    from fdtd_accelerate_amd.sim import YeeSim
    fp = {}
    if T:
        fp["rows_per_thread"] = T
    if CX:
        fp["planes_per_chunk"] = CX
    return YeeSim(*ns, npml=npml, coeff_use="e" if ce else "", engine=engine, fused_params=fp)


def _fill(sim, seed):
    rng = np.random.default_rng(seed)
    f = sim.fields
    nx, ny, nz = f.ns
    f.set_ehs(*[rng.standard_normal((nx, ny, nz)).astype(np.float32) for _ in range(6)])
    if f.ce_on:
        f.set_ces(*[(0.5 / rng.uniform(1, 6, (nx, ny, nz))).astype(np.float32) for _ in range(3)])


@pytest.mark.parametrize("ns,npml,T,CX,steps", [
    ((24, 20, 32), 0, 1, 7, 3),
    ((24, 20, 32), 0, 2, 64, 4),
    ((24, 20, 32), 0, 7, 5, 5),
    ((17, 23, 45), 0, 4, 5, 5),
    ((30, 26, 40), 4, 2, 8, 6),
    ((30, 26, 40), 4, 6, 3, 5),
    ((29, 31, 57), 5, 4, 16, 7),
    ((40, 36, 160), 6, 2, 11, 4),
    ((22, 22, 200), 3, 1, 1, 3),
    ((12, 13, 1000), 4, 3, 5, 3),
    ((26, 70, 300), 5, 7, 9, 2),
    ((20, 30, 300), 4, 5, 7, 3),
    ((26, 70, 300), 5, 11, 9, 2),
    ((30, 40, 48), 4, 15, 8, 3),
    ((17, 45, 530), 3, 13, 64, 2),
])
def test_fused_matches_inplace(ns, npml, T, CX, steps):
    """Random fields (walls included) + CE + CPML + a point and a plane source + probes:
    fused engine == in-place engine after `steps` steps, fields, psi and probe series."""
    from fdtd_accelerate_amd.sources import gaussian_pulse
    sims = [_sim(ns, npml, "kemp"), _sim(ns, npml, "fused", T, CX)]
    nx, ny, nz = ns
    tf = gaussian_pulse(5.0, 3.0, 8.0)
    for s in sims:
        _fill(s, 42)
        s.add_source("ez", (nx // 2, ny // 2, nz // 2), (nx // 2, ny // 2, nz // 2), tf)
        s.add_source("hy", (3, 1, 0), (3, ny - 2, nz - 1), tf, 0.25)
        s.add_probe("ez", (nx // 2 + 1, ny // 2, nz // 2))
        s.add_probe("hx", (1, 2, 3))
        s.add_probe("hz", (nx - 1, ny - 1, nz - 1))
    series = [s.run(steps) for s in sims]
    for name in ("ex", "ey", "ez", "hx", "hy", "hz"):
        a = sims[0].field(name).cpu().numpy()
        b = sims[1].field(name).cpu().numpy()
        assert same_values(a, b), (name, np.abs(a - b).max())
    assert same_values(series[0], series[1])
    if npml:
        for key in sims[0].cpml.psiE:
            for c in range(2):
                assert same_values(sims[0].cpml.psiE[key][c].cpu().numpy(), sims[1].cpml.psiE[key][c].cpu().numpy()), ("psiE", key, c)
                assert same_values(sims[0].cpml.psiH[key][c].cpu().numpy(), sims[1].cpml.psiH[key][c].cpu().numpy()), ("psiH", key, c)


@pytest.mark.parametrize("ns,npml,T,CX,steps", [
    ((24, 20, 32), 0, 3, 5, 3),
    ((30, 26, 40), 4, 7, 8, 5),
    ((29, 31, 57), 5, 5, 16, 4),
    ((12, 13, 1000), 4, 2, 5, 3),
    ((29, 40, 300), 5, 15, 16, 3),
])
def test_fused_ce_table_matches_inplace(ns, npml, T, CX, steps):
    """CE held as uint8 material indices + table (Fields.compress_ce) on the fused engine,
    the sweep reading one uint16 triple index per cell: equal to the in-place engine
    reading the float CE arrays, also after new CE values mid-run (indices rebuilt)."""
    from fdtd_accelerate_amd.sources import gaussian_pulse
    sims = [_sim(ns, npml, "kemp"), _sim(ns, npml, "fused", T, CX)]
    nx, ny, nz = ns
    rng = np.random.default_rng(5)
    vals = (0.5 / np.array([1.0, 1.7, 2.25, 4.0, 6.5, 9.0, 12.0])).astype(np.float32)
    mats = [rng.integers(0, len(vals), (nx, ny, nz)) for _ in range(3)]
    tf = gaussian_pulse(5.0, 3.0, 8.0)
    for s in sims:
        _fill(s, 11)
        s.fields.set_ces(*[vals[m] for m in mats])
        s.add_source("ez", (nx // 2, ny // 2, nz // 2), (nx // 2, ny // 2, nz // 2), tf)
        s.add_probe("ey", (nx // 2, ny // 2 + 1, nz // 2))
    assert sims[1].fields.compress_ce()
    series = [s.run(steps) for s in sims]
    st = sims[1]._fused
    # the plain sweep reads one uint16 triple index per cell (7 values per component)
    assert st.ce_tri is not None and st.ce_tri_lut.shape[0] <= 7 ** 3 and st.bytes_per_cell == 50.0
    for name in ("ex", "ey", "ez", "hx", "hy", "hz"):
        a = sims[0].field(name).cpu().numpy()
        b = sims[1].field(name).cpu().numpy()
        assert same_values(a, b), (name, np.abs(a - b).max())
    assert same_values(series[0], series[1])
    # new CE values mid-run: the material indices and the triples are rebuilt
    vals2 = (0.5 / np.array([1.3, 2.0, 5.0])).astype(np.float32)
    mats2 = [rng.integers(0, len(vals2), (nx, ny, nz)) for _ in range(3)]
    old = st.ce_tri
    for s in sims:
        s.fields.set_ces(*[vals2[m] for m in mats2])
    assert sims[1].fields.compress_ce()
    series = [s.run(steps) for s in sims]
    assert st.ce_tri is not None and st.ce_tri is not old and st.ce_tri_lut.shape[0] <= 4 ** 3  # pads keep 0.5
    for name in ("ex", "ey", "ez", "hx", "hy", "hz"):
        assert same_values(sims[0].field(name).cpu().numpy(), sims[1].field(name).cpu().numpy()), ("new CE", name)
    assert same_values(series[0], series[1])


def _pair_box(ns, npml):
    """The two-step sweep's box (xb, xe, yb, ye, kb, ke) or None when the grid runs single steps."""
    nx, ny, nz = ns
    p = npml
    box = (p + 2, nx - 2 - p, p + 2, ny - 2 - p, 32 * ((p + 33) // 32), 32 * ((nz - 2 - p) // 32))
    if box[1] - box[0] >= 4 and box[3] - box[2] >= 4 and box[5] > box[4]:
        return box
    return None


@pytest.mark.parametrize("ns,npml,T,CX,steps,table", [
    ((22, 22, 200), 3, 1, 1, 3, False),
    ((40, 36, 160), 6, 2, 11, 4, False),
    ((26, 70, 300), 5, 7, 9, 2, False),
    ((20, 30, 300), 4, 5, 7, 5, True),
    ((17, 45, 530), 3, 9, 64, 6, False),
    ((29, 40, 300), 5, 9, 16, 5, True),
    ((33, 27, 700), 4, 4, 3, 7, False),
    ((12, 13, 1000), 4, 3, 5, 3, False),
])
def test_fused_pairs_match_inplace(ns, npml, T, CX, steps, table):
    """Step pairs (two steps per interior sweep, shell passes into a third buffer set):
    equal to the in-place engine after odd and even step counts -- fields, psi, probe
    series -- with sources and probes in the CPML shell, on both fringes of the two-step
    box, inside it and across its z segments (256 cells)."""
    from fdtd_accelerate_amd.sources import gaussian_pulse
    from fdtd_accelerate_amd.sim import YeeSim
    nx, ny, nz = ns
    fp = {"rows_per_thread": T, "planes_per_chunk": CX, "step_pairs": True}
    sims = [_sim(ns, npml, "kemp"), YeeSim(*ns, npml=npml, coeff_use="e", engine="fused", fused_params=fp)]
    tf = gaussian_pulse(5.0, 3.0, 8.0)
    rng = np.random.default_rng(17)
    vals = (0.5 / np.array([1.0, 2.25, 4.0, 9.0])).astype(np.float32)
    mats = [rng.integers(0, len(vals), (nx, ny, nz)) for _ in range(3)]
    box = _pair_box(ns, npml)
    xb, xe, yb, ye, kb, ke = box or (npml + 2, nx - 2 - npml, npml + 2, ny - 2 - npml, 32, 64)
    kz = min(ke - 1, 256) if ke > 256 else (kb + ke) // 2
    probes = [("ez", (nx // 2 + 1, ny // 2, nz // 2)), ("hx", (1, 2, 3)), ("hz", (nx - 1, ny - 1, nz - 1)),
              ("hy", (xb, yb, kb)), ("ey", (xe - 1, ye - 1, ke - 1)), ("hx", (xe - 1, ye - 1, ke - 1)),
              ("ex", (xb + 1, yb + 1, kz)), ("hz", (xb + 1, yb + 1, kz - 1))]
    for s in sims:
        _fill(s, 23)
        if table:
            s.fields.set_ces(*[vals[m] for m in mats])
        s.add_source("ez", (nx // 2, ny // 2, nz // 2), (nx // 2, ny // 2, nz // 2), tf)
        s.add_source("hy", (3, 1, 0), (3, ny - 2, nz - 1), tf, 0.25)
        s.add_source("ey", (1, 1, kz), (nx - 2, ny - 2, kz), tf, 0.5)
        s.add_source("hx", (nx // 2, yb + 1, 0), (nx // 2, yb + 1, nz - 1), tf, 0.125)
        for comp, pt in probes:
            s.add_probe(comp, pt)
    if table:
        assert sims[1].fields.compress_ce()
    series = [s.run(steps) for s in sims]
    assert sims[1]._fused.pairs
    for name in ("ex", "ey", "ez", "hx", "hy", "hz"):
        a = sims[0].field(name).cpu().numpy()
        b = sims[1].field(name).cpu().numpy()
        assert same_values(a, b), (name, np.abs(a - b).max(), np.argwhere(canon_arr(a) != canon_arr(b))[:4])
    assert same_values(series[0], series[1]), np.argwhere(series[0] != series[1])[:8]
    for key in sims[0].cpml.psiE:
        for c in range(2):
            assert same_values(sims[0].cpml.psiE[key][c].cpu().numpy(), sims[1].cpml.psiE[key][c].cpu().numpy()), ("psiE", key, c)
            assert same_values(sims[0].cpml.psiH[key][c].cpu().numpy(), sims[1].cpml.psiH[key][c].cpu().numpy()), ("psiH", key, c)
    # continuing with another odd call keeps equality (parity of the buffer sets)
    series = [s.run(3) for s in sims]
    assert same_values(series[0], series[1])
    for name in ("ex", "hz"):
        assert same_values(sims[0].field(name).cpu().numpy(), sims[1].field(name).cpu().numpy()), name


def canon_arr(a):
    return np.ascontiguousarray(a, dtype=np.float32) + np.float32(0.0)


def test_fused_field_arena(monkeypatch):
    """Field sets carved from one allocation with staggered starts (the layout used for
    arrays of 1 GiB and more, forced here on a small grid, for the in-place engine's six
    arrays and the fused engine's twelve): the fused engine equals the in-place engine
    over odd and even step counts, and its twelve arrays sit in one storage at
    k * (size + gap) bytes, gap = size/8 rounded up to 2 MiB."""
    from fdtd_accelerate_amd import fields as F
    from fdtd_accelerate_amd.sources import gaussian_pulse
    monkeypatch.setattr(F, "ARENA_MIN_BYTES", 1)
    ns, npml = (30, 26, 40), 4
    nx, ny, nz = ns
    sims = [_sim(ns, npml, "kemp"), _sim(ns, npml, "fused", 3, 8)]
    tf = gaussian_pulse(5.0, 3.0, 8.0)
    for s in sims:
        _fill(s, 7)
        s.add_source("ez", (nx // 2, ny // 2, nz // 2), (nx // 2, ny // 2, nz // 2), tf)
        s.add_probe("hy", (nx // 2, ny // 2, nz // 2 + 1))
    for steps in (3, 4):
        series = [s.run(steps) for s in sims]
        assert same_values(series[0], series[1])
        for name in ("ex", "ey", "ez", "hx", "hy", "hz"):
            a = sims[0].field(name).cpu().numpy()
            b = sims[1].field(name).cpu().numpy()
            assert same_values(a, b), (steps, name, np.abs(a - b).max())
    ref = sims[0].fields.ehs
    assert all(t.untyped_storage().data_ptr() == ref[0].untyped_storage().data_ptr() for t in ref)
    st = sims[1]._fused
    arrays = sims[1].fields.ehs + st.alt
    base = arrays[0].untyped_storage().data_ptr()
    assert all(t.untyped_storage().data_ptr() == base for t in arrays)
    nbytes = arrays[0].numel() * 4
    step = nbytes + -(-(nbytes // 8) // (2 << 20)) * (2 << 20)
    assert sorted(t.data_ptr() - base for t in arrays) == [k * step for k in range(12)]


def test_fused_continues_across_calls_and_engines():
    """Steps split over several run() calls (odd/even parity, buffer swaps) and an engine
    switch mid-run equal one continuous run."""
    from fdtd_accelerate_amd.sources import gaussian_pulse
    ns, npml = (26, 28, 36), 4
    tf = gaussian_pulse(4.0, 2.0, 8.0)
    a = _sim(ns, npml, "kemp")
    b = _sim(ns, npml, "fused", 2, 6)
    for s in (a, b):
        _fill(s, 7)
        s.add_source("ex", (13, 14, 18), (13, 14, 18), tf)
    a.run(9)
    b.run(3)
    b.run(1)
    b.engine = "kemp"
    b.run(2)
    b.engine = "fused"
    b.run(3)
    for name in ("ex", "ey", "ez", "hx", "hy", "hz"):
        assert same_values(a.field(name).cpu().numpy(), b.field(name).cpu().numpy()), name


def test_instances_added_mid_run():
    """A Drude box added after the whole-step engines have run rebuilds them (the state
    carries over): fused and slab equal the in-place engine, and a slab run with the box
    is refused."""
    from fdtd_accelerate_amd import Drude
    from fdtd_accelerate_amd.sources import gaussian_pulse
    ns, npml = (26, 28, 72), 4
    tf = gaussian_pulse(4.0, 2.0, 8.0)
    a = _sim(ns, npml, "kemp")
    b = _sim(ns, npml, "fused", 3, 8)
    for s in (a, b):
        _fill(s, 21)
        s.add_source("ez", (13, 14, 30), (13, 14, 30), tf)
        s.run(3)
    first = b._fused
    for s in (a, b):
        Drude(s.fields, (6, 6, 20), (18, 20, 50), 0.5, 0.05)
        s.run(4)
    assert b._fused is not first and b._fused.drude is not None
    for name in ("ex", "ey", "ez", "hx", "hy", "hz"):
        assert same_values(a.field(name).cpu().numpy(), b.field(name).cpu().numpy()), name
    c = _slab(ns, npml, 4)
    _fill(c, 21)
    c.run(2)
    Drude(c.fields, (6, 6, 20), (18, 20, 50), 0.5, 0.05)
    with pytest.raises(ValueError, match="Drude"):
        c.run(1)


@pytest.mark.parametrize("z_phase,drude", [(None, False), (0, False), (8, False), (16, False), (28, False),
                                            (16, True), (None, True)])
def test_fused_line_phases_match_inplace(z_phase, drude):
    """Field arrays placed at each 128-byte line phase (FusedStepper z_phase; None: the
    automatic choice, 0 with a Drude box) move the sweep's z range and the segment
    origin, not the results: fields, psi and probes equal the in-place engine, with and
    without a Drude box crossing segment boundaries."""
    from fdtd_accelerate_amd import Drude
    from fdtd_accelerate_amd.fused import FusedStepper, z_line_phase
    from fdtd_accelerate_amd.sources import gaussian_pulse
    ns, npml, steps = (24, 30, 416), 6, 5
    nx, ny, nz = ns
    sims = [_sim(ns, npml, "kemp"), _sim(ns, npml, "fused")]
    tf = gaussian_pulse(5.0, 3.0, 8.0)
    rng = np.random.default_rng(9)
    vals = (0.5 / np.array([1.0, 2.0, 3.5, 7.0, 11.0])).astype(np.float32)
    mats = [rng.integers(0, len(vals), ns) for _ in range(3)]
    for s in sims:
        _fill(s, 17)
        s.fields.set_ces(*[vals[m] for m in mats])
        s.add_source("ez", (nx // 2, ny // 2, nz // 2), (nx // 2, ny // 2, nz // 2), tf)
        s.add_probe("ey", (nx // 2, ny // 2, 20))
        s.add_probe("hx", (9, 9, 405))
        if drude:  # across the shell and the 256-cell segment boundaries
            Drude(s.fields, (5, 5, 20), (18, 24, 300), 0.5, 0.05)
    assert sims[1].fields.compress_ce()
    st = FusedStepper(sims[1], rows_per_thread=5, planes_per_chunk=8, z_phase=z_phase)
    sims[1]._fused = st
    want = (0 if drude else z_line_phase(nz, sims[1].fields.nz_pitch, npml)) if z_phase is None else z_phase
    assert st.z_phase == want
    box = st.sweep_box()
    assert (box[4] + st.z_phase) % 32 == 0 and box[4] >= npml + 1 and box[5] <= nz - 1 - npml
    series = [s.run(steps) for s in sims]
    for name in ("ex", "ey", "ez", "hx", "hy", "hz"):
        a = sims[0].field(name).cpu().numpy()
        b = sims[1].field(name).cpu().numpy()
        assert same_values(a, b), (name, np.abs(a - b).max())
    assert same_values(series[0], series[1])
    for key in sims[0].cpml.psiE:
        for c in range(2):
            assert same_values(sims[0].cpml.psiE[key][c].cpu().numpy(), sims[1].cpml.psiE[key][c].cpu().numpy())
            assert same_values(sims[0].cpml.psiH[key][c].cpu().numpy(), sims[1].cpml.psiH[key][c].cpu().numpy())


def test_fused_dipole128_vs_oracle():
    """The 128^3 vacuum dipole run (CPML 10, point source, probe) on the fused engine, 400 steps:
    probe series and fields equal the float32 oracle."""
    from fdtd_accelerate_amd.sim import YeeSim
    from fdtd_accelerate_amd.sources import gaussian_pulse
    steps = 400
    sim = YeeSim(128, 128, 128, npml=10, coeff_use="", engine="fused")
    sim.add_source("ez", (64, 64, 64), (64, 64, 64), gaussian_pulse())
    sim.add_probe("ez", (96, 64, 64))
    series = sim.run(steps)[:, 0]
    o = O.OracleSim(128, 128, 128, 0, 10, np.float32)
    o.add_source("ez", (64, 64, 64), (64, 64, 64), O.pulse)
    o.add_probe("ez", (96, 64, 64))
    o.run(steps)
    assert same_values(series, np.array(o.series, np.float32)[:, 0])
    for name in ("ex", "ey", "ez", "hx", "hy", "hz"):
        assert same_values(sim.field(name).cpu().numpy(), o.field(name)), name


def test_fused_rejects_unsupported():
    from fdtd_accelerate_amd.fused import FusedStepper
    from fdtd_accelerate_amd.sim import YeeSim
    s = YeeSim(24, 24, 24, npml=4, cpml_faces=("fb", "f", "fb"), engine="fused")
    with pytest.raises(ValueError):
        FusedStepper(s)
    s2 = YeeSim(8, 8, 16, coeff_use="eh", engine="fused")
    with pytest.raises(ValueError):
        FusedStepper(s2)


def _slab(ns, npml, S, ce=True):
    from fdtd_accelerate_amd.sim import YeeSim
    return YeeSim(*ns, npml=npml, coeff_use="e" if ce else "", engine="slab", fused_params={"slab_planes": S})


@pytest.mark.parametrize("ns,npml,S,steps,ce", [
    ((24, 20, 32), 0, 1, 3, True),
    ((24, 20, 32), 0, 5, 4, False),
    ((30, 26, 40), 4, 1, 5, True),
    ((30, 26, 40), 4, 7, 6, True),
    ((29, 31, 57), 5, 4, 7, True),
    ((40, 36, 160), 6, 3, 4, False),
    ((12, 13, 1000), 4, 2, 3, True),
    ((26, 70, 300), 5, 26, 2, True),
    ((33, 64, 128), 8, 64, 3, True),
])
def test_slab_matches_inplace(ns, npml, S, steps, ce):
    """Slab engine == in-place engine byte for byte (signed zeros included): fields, psi
    and probe series, with a point, a plane (spatial array) and an overwrite source."""
    from fdtd_accelerate_amd.sources import gaussian_pulse
    sims = [_sim(ns, npml, "kemp", ce=ce), _slab(ns, npml, S, ce)]
    nx, ny, nz = ns
    tf = gaussian_pulse(5.0, 3.0, 8.0)
    sv = np.random.default_rng(3).standard_normal((ny - 2, nz)).astype(np.float32)
    for s in sims:
        _fill(s, 42)
        s.add_source("ez", (nx // 2, ny // 2, nz // 2), (nx // 2, ny // 2, nz // 2), tf)
        s.add_source("hy", (3, 1, 0), (3, ny - 2, nz - 1), tf, sv)
        s.add_source("ex", (nx - 1, ny - 1, 0), (nx - 1, ny - 1, nz - 1), tf, 0.25, True)
        s.add_probe("ez", (nx // 2 + 1, ny // 2, nz // 2))
        s.add_probe("hx", (1, 2, 3))
        s.add_probe("hz", (nx - 1, ny - 1, nz - 1))
        s.add_probe("ey", (0, 0, 0))
    series = [s.run(steps) for s in sims]
    for name in ("ex", "ey", "ez", "hx", "hy", "hz"):
        a = sims[0].field(name).cpu().numpy()
        b = sims[1].field(name).cpu().numpy()
        assert a.tobytes() == b.tobytes(), (name, np.abs(a - b).max())
    assert series[0].tobytes() == series[1].tobytes()
    if npml:
        for key in sims[0].cpml.psiE:
            for c in range(2):
                assert sims[0].cpml.psiE[key][c].cpu().numpy().tobytes() == sims[1].cpml.psiE[key][c].cpu().numpy().tobytes(), ("psiE", key, c)
                assert sims[0].cpml.psiH[key][c].cpu().numpy().tobytes() == sims[1].cpml.psiH[key][c].cpu().numpy().tobytes(), ("psiH", key, c)


@pytest.mark.parametrize("ns,npml,S,steps", [
    ((24, 20, 32), 0, 5, 4),
    ((30, 26, 40), 4, 1, 5),
    ((29, 31, 57), 5, 4, 7),
    ((26, 70, 300), 5, 26, 2),
])
def test_slab_ce_table_matches_inplace(ns, npml, S, steps):
    """CE held as uint8 material indices + table (Fields.compress_ce) on the slab engine:
    byte-equal to the in-place engine reading the float CE arrays (fields, psi, probes)."""
    from fdtd_accelerate_amd.sources import gaussian_pulse
    sims = [_sim(ns, npml, "kemp"), _slab(ns, npml, S)]
    nx, ny, nz = ns
    rng = np.random.default_rng(8)
    vals = (0.5 / np.array([1.0, 1.7, 2.25, 4.0, 6.5, 9.0, 12.0])).astype(np.float32)
    mats = [rng.integers(0, len(vals), (nx, ny, nz)) for _ in range(3)]
    tf = gaussian_pulse(5.0, 3.0, 8.0)
    for s in sims:
        _fill(s, 13)
        s.fields.set_ces(*[vals[m] for m in mats])
        s.add_source("ez", (nx // 2, ny // 2, nz // 2), (nx // 2, ny // 2, nz // 2), tf)
        s.add_probe("ey", (nx // 2, ny // 2 + 1, nz // 2))
        s.add_probe("ex", (1, 1, 1))
    assert sims[1].fields.compress_ce()
    series = [s.run(steps) for s in sims]
    for name in ("ex", "ey", "ez", "hx", "hy", "hz"):
        a = sims[0].field(name).cpu().numpy()
        b = sims[1].field(name).cpu().numpy()
        assert a.tobytes() == b.tobytes(), (name, np.abs(a - b).max())
    assert series[0].tobytes() == series[1].tobytes()
    if npml:
        for key in sims[0].cpml.psiE:
            for c in range(2):
                assert sims[0].cpml.psiE[key][c].cpu().numpy().tobytes() == sims[1].cpml.psiE[key][c].cpu().numpy().tobytes(), ("psiE", key, c)


def test_slab_dipole128_vs_oracle():
    """The 128^3 vacuum dipole run (CPML 10, point source, probe) on the slab engine, 400
    steps: probe series and fields byte-equal to the float32 oracle."""
    from fdtd_accelerate_amd.sim import YeeSim
    from fdtd_accelerate_amd.sources import gaussian_pulse
    steps = 400
    sim = YeeSim(128, 128, 128, npml=10, coeff_use="", engine="slab")
    sim.add_source("ez", (64, 64, 64), (64, 64, 64), gaussian_pulse())
    sim.add_probe("ez", (96, 64, 64))
    series = sim.run(steps)[:, 0]
    o = O.OracleSim(128, 128, 128, 0, 10, np.float32)
    o.add_source("ez", (64, 64, 64), (64, 64, 64), O.pulse)
    o.add_probe("ez", (96, 64, 64))
    o.run(steps)
    assert series.tobytes() == np.array(o.series, np.float32)[:, 0].tobytes()
    for name in ("ex", "ey", "ez", "hx", "hy", "hz"):
        assert sim.field(name).cpu().numpy().tobytes() == o.field(name).tobytes(), name


def test_slab_engine_switch():
    """kemp -> slab -> fused -> slab mid-run equals one continuous in-place run."""
    from fdtd_accelerate_amd.sources import gaussian_pulse
    ns, npml = (26, 28, 36), 4
    tf = gaussian_pulse(4.0, 2.0, 8.0)
    a = _sim(ns, npml, "kemp")
    b = _slab(ns, npml, 5)
    for s in (a, b):
        _fill(s, 9)
        s.add_source("ez", (13, 14, 18), (13, 14, 18), tf)
        s.add_probe("ez", (14, 14, 18))
    a.run(10)
    b.engine = "kemp"
    b.run(2)
    b.engine = "slab"
    b.run(3)
    b.engine = "fused"
    b._fused_params = {"rows_per_thread": 2, "planes_per_chunk": 6}
    b.run(3)
    b.engine = "slab"
    b.run(2)
    for name in ("ex", "ey", "ez", "hx", "hy", "hz"):
        assert same_values(a.field(name).cpu().numpy(), b.field(name).cpu().numpy()), name
    assert same_values(a.series(), b.series())
