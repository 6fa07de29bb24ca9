"""TFSF plane wave on the GPU (``Tfsf``, ``mifdtd_tfsf_line`` / ``mifdtd_tfsf_faces``)
against the float32 oracle: every field byte-equal (-0 == +0) after the steps, for each
propagation axis, two polarization angles, all or some faces, with CPML and a
dielectric outside the box, on the in-place and the fused engine; and the
scattered-field invariant on the GPU run."""
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


def _pair(ns, npml, box0, box1, apply, axis, angle, wavelength, ce=None, seg=64, engine="kemp", fp=None):
    from fdtd_accelerate_amd import Tfsf
    from fdtd_accelerate_amd.sim import YeeSim
    sim = YeeSim(*ns, npml=npml, coeff_use="e" if ce is not None else "", segment_nbytes=seg, engine=engine,
                 fused_params=fp)
    f = sim.fields
    if ce is not None:
        f.set_ces(*ce)
    t = Tfsf(f, box0, box1, apply, wavelength, ("normal", axis), angle)
    pad = f.pad
    ces = None if ce is None else [np.pad(c, ((0, 0), (0, 0), (0, pad)), constant_values=np.float32(0.5)) for c in ce]
    o = O.OracleSim(*ns, pad, npml, np.float32, ces=ces)
    o.tfsf.append(O.OracleTfsf(box0, box1, apply, t.wdt, axis, angle))
    return sim, o


@pytest.mark.parametrize("axis,angle,apply,npml,seg", [
    ("x", 0.0, ("fb", "fb", "fb"), 5, 64),
    ("y", 0.7, ("fb", "fb", "fb"), 5, 64),
    ("z", 1.1, ("fb", "fb", "fb"), 0, 16),
    ("x", 0.4, ("f", "fb", "b"), 4, 64),
    ("z", 0.0, ("fb", "", "f"), 5, 32),
])
def test_tfsf_matches_oracle(axis, angle, apply, npml, seg):
    ns = (30, 28, 33)
    rng = np.random.default_rng(5)
    ce = [np.full(ns, 0.5, np.float32) for _ in range(3)]
    for c in ce:  # a dielectric slab outside the box, on the far side along each axis
        c[:, :, 26:29] = np.float32(0.5 / 3.0)
        c[24:27] = (0.5 / rng.uniform(1, 4, c[24:27].shape)).astype(np.float32)
    sim, o = _pair(ns, npml, (8, 7, 9), (21, 20, 22), apply, axis, angle, 10.0, ce, seg)
    sim.run(70)
    o.run(70)
    for k, name in enumerate(NAMES):
        assert _same(sim.field(name).cpu().numpy(), o.ehs[k]), name
    assert np.abs(o.ehs[2]).max() > 0.1 or np.abs(o.ehs[1]).max() > 0.1


@pytest.mark.parametrize("axis,angle,apply,npml,fp,table", [
    ("x", 0.0, ("fb", "fb", "fb"), 5, {"rows_per_thread": 3, "planes_per_chunk": 5}, False),
    ("y", 0.7, ("fb", "fb", "fb"), 5, {"rows_per_thread": 9, "planes_per_chunk": 64}, True),
    ("z", 1.1, ("fb", "fb", "fb"), 0, {"rows_per_thread": 2, "planes_per_chunk": 7}, False),
    ("x", 0.4, ("f", "fb", "b"), 4, None, True),
    ("z", 0.0, ("fb", "", "f"), 5, {"rows_per_thread": 15, "planes_per_chunk": 16}, False),
], ids=["x", "y_table_T9", "z_no_cpml", "x_some_faces_table", "z_T15"])
def test_tfsf_fused_matches_oracle(axis, angle, apply, npml, fp, table):
    """The plane wave on the fused engine: the sweep box is z 32..96 of nz = 132 (32..128
    without CPML), so the E and H patches of the box faces fall in the sweep (owned and
    recomputed halo rows and seed planes) and in both shell passes; the incident line
    advances around each step.  Byte-equal to the oracle over 70 steps, split into an odd
    and an even call."""
    ns = (30, 28, 132)
    rng = np.random.default_rng(5)
    if table:
        vals = (0.5 / np.array([1.0, 2.0, 3.5])).astype(np.float32)
        ce = [vals[rng.integers(0, 3, ns)] for _ in range(3)]
    else:
        ce = [np.full(ns, 0.5, np.float32) for _ in range(3)]
        for c in ce:
            c[:, :, 100:110] = np.float32(0.5 / 3.0)
            c[24:27] = (0.5 / rng.uniform(1, 4, c[24:27].shape)).astype(np.float32)
    sim, o = _pair(ns, npml, (8, 7, 20), (21, 20, 90), apply, axis, angle, 10.0, ce, 64, "fused", fp)
    if table:
        assert sim.fields.compress_ce()
    sim.run(35)
    sim.run(35)
    assert sim._fused.tfsf is not None
    o.run(70)
    for k, name in enumerate(NAMES):
        assert _same(sim.field(name).cpu().numpy(), o.ehs[k]), name
    assert np.abs(o.ehs[2]).max() > 0.1 or np.abs(o.ehs[1]).max() > 0.1


@pytest.mark.parametrize("ns,fp,rows", [
    ((30, 28, 128), {"rows_per_thread": 3, "planes_per_chunk": 8}, 3),
    ((30, 28, 128), {"rows_per_thread": 13, "planes_per_chunk": 16}, 13),
    ((24, 110, 128), None, 13),
], ids=["T3", "T13", "auto_T13"])
def test_tfsf_fused_line_phase_and_ce_triples(ns, fp, rows):
    """The plane wave on the fused engine as the 1024^3 plane-wave workload runs it: field
    arrays at a nonzero line phase (nz_pitch 128), the sweep reading triple CE indices,
    13 rows per block (the 1024-thread build; the automatic choice on a grid at least 104
    rows tall); byte-equal to the oracle over 40 steps."""
    rng = np.random.default_rng(17)
    vals = (0.5 / np.array([1.0, 2.0, 3.5, 8.0])).astype(np.float32)
    ce = [vals[rng.integers(0, len(vals), ns)] for _ in range(3)]
    sim, o = _pair(ns, 5, (8, 7, 20), (ns[0] - 9, ns[1] - 8, 100), ("fb", "fb", "fb"), "x", 0.3, 10.0, ce, 64,
                   "fused", fp)
    assert sim.fields.compress_ce()
    sim.run(17)
    sim.run(23)
    st = sim._fused
    assert st.tfsf is not None and st.z_phase != 0 and st.ce_tri is not None and st.T == rows
    o.run(40)
    for k, name in enumerate(NAMES):
        assert _same(sim.field(name).cpu().numpy(), o.ehs[k]), name
    assert np.abs(o.ehs[2]).max() > 0.1 or np.abs(o.ehs[1]).max() > 0.1


def test_tfsf_fused_shell_only_matches_oracle():
    """A grid with no sweep box (nz = 33 with CPML 5): every patch is applied by the shell
    passes."""
    ns = (30, 28, 33)
    sim, o = _pair(ns, 5, (8, 7, 9), (21, 20, 22), ("fb", "fb", "fb"), "y", 0.3, 10.0, engine="fused")
    sim.run(70)
    o.run(70)
    for k, name in enumerate(NAMES):
        assert _same(sim.field(name).cpu().numpy(), o.ehs[k]), name


def test_tfsf_scattered_field_stays_quiet_on_gpu():
    from fdtd_accelerate_amd import Tfsf
    from fdtd_accelerate_amd.sim import YeeSim
    n, npml, lo, hi = 48, 8, 14, 33
    sim = YeeSim(n, n, n, npml=npml, coeff_use="")
    Tfsf(sim.fields, (lo,) * 3, (hi,) * 3, ("fb", "fb", "fb"), 12.0, ("normal", "y"), 0.3)
    sim.run(120)
    f = [sim.field(nm).cpu().numpy() for nm in NAMES]
    inside = max(np.abs(x[lo + 1:hi, lo + 1:hi, lo + 1:hi]).max() for x in f)
    m = np.zeros((n, n, n), bool)
    m[npml + 1:n - npml - 1, npml + 1:n - npml - 1, npml + 1:n - npml - 1] = True
    m[lo - 2:hi + 3, lo - 2:hi + 3, lo - 2:hi + 3] = False
    outside = max(np.abs(x[m]).max() for x in f)
    assert inside > 0.5 and outside < 1e-5 * inside, (inside, outside)


def test_tfsf_arguments_and_engines():
    from fdtd_accelerate_amd import Tfsf
    from fdtd_accelerate_amd.batch import BatchedYee
    from fdtd_accelerate_amd.sim import YeeSim
    s = YeeSim(16, 16, 16, npml=0, coeff_use="")
    with pytest.raises(NotImplementedError):
        Tfsf(s.fields, (3, 3, 3), (12, 12, 12), ("fb",) * 3, 8.0, ("oblique", 0.3, 0.2))
    with pytest.raises(ValueError):
        Tfsf(s.fields, (0, 3, 3), (12, 12, 12), ("fb",) * 3, 8.0)
    t = Tfsf(s.fields, (None, None, None), (None, None, None), ("fb",) * 3, 8.0)
    assert t.pt0 == [1, 1, 1] and t.pt1 == [15, 15, 15]
    with pytest.raises(ValueError, match="TFSF"):
        BatchedYee([s])
    s2 = YeeSim(16, 16, 16, npml=0, coeff_use="", engine="slab")
    Tfsf(s2.fields, (3, 3, 3), (12, 12, 12), ("fb",) * 3, 8.0)
    with pytest.raises(ValueError, match="TFSF"):
        s2.run(1)
    # fused: one plane wave, no step pairs with it, no source on a patch cell of its component
    s3 = YeeSim(16, 16, 64, npml=4, coeff_use="", engine="fused", fused_params={"step_pairs": True})
    Tfsf(s3.fields, (3, 3, 3), (12, 12, 60), ("fb",) * 3, 8.0)
    with pytest.raises(ValueError, match="TFSF"):
        s3.run(1)
    s4 = YeeSim(16, 16, 64, npml=4, coeff_use="", engine="fused")
    t4 = Tfsf(s4.fields, (3, 3, 3), (12, 12, 60), ("fb",) * 3, 8.0, ("normal", "z"))
    P = t4.patches_e
    s4.add_source(("ex", "ey", "ez")[P.comp[0]], (P.box[0][0], P.box[0][2], P.box[0][4]),
                  (P.box[0][0], P.box[0][2], P.box[0][4]), lambda t: 1.0)
    with pytest.raises(ValueError, match="TFSF"):
        s4.run(1)


def test_tfsf_plane_wave_on_drude_sphere_matches_oracle():
    """A Drude sphere (per-cell plasma frequency inside a radius, 0 outside) in a TFSF
    box with CPML around it: the scatterer case of both features together, byte-equal
    to the oracle (Drude right after the core E update, the plane wave after the
    sources, as their priorities order them)."""
    from fdtd_accelerate_amd import Drude, Tfsf
    from fdtd_accelerate_amd.sim import YeeSim
    ns, npml, steps = (34, 32, 30), 5, 80
    box0, box1 = (9, 8, 8), (24, 23, 21)
    d0, d1 = (13, 12, 11), (20, 19, 18)
    shape = tuple(b - a + 1 for a, b in zip(d0, d1))
    g = np.indices(shape).astype(np.float32)
    c = np.array([(s - 1) / 2 for s in shape], np.float32).reshape(3, 1, 1, 1)
    inside = ((g - c) ** 2).sum(0) <= 3.2 ** 2
    wp = np.where(inside, np.float32(0.9), np.float32(0)).astype(np.float32)
    sim = YeeSim(*ns, npml=npml, coeff_use="")
    f = sim.fields
    t = Tfsf(f, box0, box1, ("fb", "fb", "fb"), 9.0, ("normal", "z"), 0.5)
    dr = Drude(f, d0, d1, wp, 0.05)
    o = O.OracleSim(*ns, f.pad, npml, np.float32)
    o.tfsf.append(O.OracleTfsf(box0, box1, ("fb", "fb", "fb"), t.wdt, "z", 0.5))
    cja, cjb = O.drude_coefficients(wp, np.full(shape, 0.05, np.float32), f.dt)
    od = O.OracleDrude(d0, d1, [cja] * 3, [cjb] * 3)
    o.drude.append(od)
    sim.run(steps)
    o.run(steps)
    for k, name in enumerate(NAMES):
        assert _same(sim.field(name).cpu().numpy(), o.ehs[k]), name
    for c_ in range(3):
        assert _same(dr.J[c_].cpu().numpy(), od.J[c_]), ("J", c_)
    assert max(np.abs(x).max() for x in od.J) > 1e-3


def test_tfsf_plane_wave_on_drude_sphere_fused_matches_oracle():
    """The Drude sphere in a TFSF box on the fused engine (both features in its sweep and
    shell passes; the sphere and box faces inside the sweep box z 32..64), byte-equal to
    the oracle, auxiliary current included."""
    from fdtd_accelerate_amd import Drude, Tfsf
    from fdtd_accelerate_amd.sim import YeeSim
    ns, npml, steps = (34, 32, 100), 5, 80
    box0, box1 = (9, 8, 20), (24, 23, 80)
    d0, d1 = (13, 12, 40), (20, 19, 47)
    shape = tuple(b - a + 1 for a, b in zip(d0, d1))
    g = np.indices(shape).astype(np.float32)
    c = np.array([(s - 1) / 2 for s in shape], np.float32).reshape(3, 1, 1, 1)
    inside = ((g - c) ** 2).sum(0) <= 3.2 ** 2
    wp = np.where(inside, np.float32(0.9), np.float32(0)).astype(np.float32)
    sim = YeeSim(*ns, npml=npml, coeff_use="", engine="fused")
    f = sim.fields
    t = Tfsf(f, box0, box1, ("fb", "fb", "fb"), 9.0, ("normal", "x"), 0.5)
    dr = Drude(f, d0, d1, wp, 0.05)
    o = O.OracleSim(*ns, f.pad, npml, np.float32)
    o.tfsf.append(O.OracleTfsf(box0, box1, ("fb", "fb", "fb"), t.wdt, "x", 0.5))
    cja, cjb = O.drude_coefficients(wp, np.full(shape, 0.05, np.float32), f.dt)
    od = O.OracleDrude(d0, d1, [cja] * 3, [cjb] * 3)
    o.drude.append(od)
    sim.run(steps // 2 + 1)
    sim.run(steps // 2 - 1)
    assert sim._fused.tfsf is t and sim._fused.drude is dr
    o.run(steps)
    for k, name in enumerate(NAMES):
        assert _same(sim.field(name).cpu().numpy(), o.ehs[k]), name
    for c_ in range(3):
        assert _same(dr.J[c_].cpu().numpy(), od.J[c_]), ("J", c_)
    assert max(np.abs(x).max() for x in od.J) > 1e-3


@pytest.mark.parametrize("axis,periodic,apply,angle", [
    ("x", "yz", ("fb", "", ""), 0.6),
    ("z", "x", ("", "fb", "fb"), 1.2),
])
def test_tfsf_on_periodic_grid_matches_oracle(axis, periodic, apply, angle):
    """A plane wave in a grid periodic across the propagation axis (CPML only on the
    other faces), a dielectric block inside the box: the periodic copy runs after the
    incident correction of the same half step, as the priorities order them ('incident'
    before 'pbc'), byte-equal to the oracle."""
    from fdtd_accelerate_amd import Pbc, Tfsf
    from fdtd_accelerate_amd.sim import YeeSim
    ns, npml, steps = (30, 26, 28), 5, 70
    faces = tuple("" if "xyz"[a] in periodic else "fb" for a in range(3))
    # box: the full periodic extent (1 .. n-2), 8 cells of margin elsewhere
    box0 = tuple(1 if "xyz"[a] in periodic else 8 for a in range(3))
    box1 = tuple(ns[a] - 2 if "xyz"[a] in periodic else ns[a] - 9 for a in range(3))
    ce = [np.full(ns, 0.5, np.float32) for _ in range(3)]
    for c in ce:
        c[12:17, 10:15, 11:16] = np.float32(0.5 / 4.0)
    sim = YeeSim(*ns, npml=npml, coeff_use="e", cpml_faces=faces)
    f = sim.fields
    f.set_ces(*ce)
    t = Tfsf(f, box0, box1, apply, 9.0, ("normal", axis), angle)
    Pbc(f, periodic)
    pad = f.pad
    o = O.OracleSim(*ns, pad, npml, np.float32,
                    ces=[np.pad(c, ((0, 0), (0, 0), (0, pad)), constant_values=np.float32(0.5)) for c in ce],
                    cpml_faces=tuple((("xyz"[a] not in periodic),) * 2 for a in range(3)), pbc=periodic)
    o.tfsf.append(O.OracleTfsf(box0, box1, apply, t.wdt, axis, angle))
    sim.run(steps)
    o.run(steps)
    for k, name in enumerate(NAMES):
        assert _same(sim.field(name).cpu().numpy(), o.ehs[k]), name
    assert max(np.abs(x).max() for x in o.ehs) > 0.1
