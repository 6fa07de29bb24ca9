"""CPU tests that pin the oracle (oracle/) before it is trusted as the GPU checker.

* core update: bit-exact against golden vectors of the reference's own compiled KEMP
  core (tests/golden/kemp_core.npz), against the NumPy slice form of K7/naive/core.py,
  and (where /root/reference exists) against the live reference build;
* CPML: closed-form tables, a = 0 is a no-op, exact 1-D reduction against an
  independent position-based 1-D CPML, x/y/z cyclic symmetry, absorption (energy
  decay) and a 1-D reflection bound;
* sources: the pulse's closed form and the tstep-starts-at-1 rule.
"""
import hashlib
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import golden_cases, split_cases, split_expected


# ------------------------------------------------------------------ core vs golden
def _oracle_run(meta, ehs, ces, chs):
    a = [x.copy() for x in ehs]
    for _ in range(meta["tmax"]):
        if meta["ufunc"] in ("e", "eh"):
            O.update_e(a, ces, meta["pad"])
        if meta["ufunc"] in ("h", "eh"):
            O.update_h(a, chs, meta["pad"])
    return a


def test_oracle_matches_reference_golden(golden):
    n = 0
    for meta, ehs, ces, chs, out in golden_cases(golden):
        got = _oracle_run(meta, ehs, ces, chs)
        for c, (g, r) in enumerate(zip(got, out)):
            assert np.array_equal(g, r), (meta, c, np.abs(g - r).max())
        n += 1
    assert n == int(golden["ncases"]) == 20


def test_oracle_matches_reference_golden_double():
    """The float64 oracle against the reference's 'double' programs (pads 0 and 1)."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "kemp_core_f64.npz"))
    n = 0
    for meta, ehs, ces, chs, out in golden_cases(g):
        assert ehs[0].dtype == np.float64
        got = _oracle_run(meta, ehs, ces, chs)
        for c, (x, r) in enumerate(zip(got, out)):
            assert np.array_equal(x, r) and x.tobytes() == r.tobytes(), (meta, c)
        n += 1
    assert n == int(g["ncases"]) == 20


def test_oracle_matches_reference_split_golden():
    """The reference's split programs (K7/cpu/src/core_split.c over element ranges: the
    whole grid, core_split.py's 'pre' / 'post' ranges, a range ending off a vector
    boundary, the last vector, an empty range; float pads 3 and 0, double pad 1) equal
    the oracle's whole-grid update restricted to the vectors that start in the range."""
    n = 0
    for meta, ehs, coef, out in split_cases():
        want = split_expected(meta, ehs, coef)
        for c, (w, r) in enumerate(zip(want, out)):
            assert w.tobytes() == r.tobytes(), (meta, c)
        n += 1
    assert n == 48


def test_oracle_matches_reference_digest(golden):
    from tests.golden.make_golden import big_inputs
    ehs, ces = big_inputs(golden["big_seed"])
    for _ in range(10):
        O.update_e(ehs, ces, 0)
    h = hashlib.sha256(b"".join(o.tobytes() for o in ehs[:3])).hexdigest()
    assert h == str(golden["big_sha256"])


@pytest.mark.parametrize("nz,pad", [(16, 0), (13, 3), (30, 2), (61, 3)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_oracle_matches_naive_slices(nz, pad, dtype):
    rng = np.random.default_rng(nz * 7 + pad)
    nx, ny, nzp = 6, 7, nz + pad
    mk = lambda: np.pad(rng.random((nx, ny, nz)).astype(dtype), ((0, 0), (0, 0), (0, pad)))
    ehs = [mk() for _ in range(6)]
    ces = [mk() for _ in range(3)]
    chs = [mk() for _ in range(3)]
    a = [x.copy() for x in ehs]
    b = [x.copy() for x in ehs]
    for _ in range(4):
        O.update_e(a, ces, pad)
        O.update_h(a, chs, pad)
        O.naive_update_e(b, ces, pad)
        O.naive_update_h(b, chs, pad)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.skipif(not O.ref_available(), reason="reference sources not present (GPU box)")
@pytest.mark.parametrize("dtype,nz", [(np.float32, 61), (np.float32, 64), (np.float64, 63)])
def test_oracle_matches_live_reference(dtype, nz):
    align = 4 if dtype == np.float32 else 2
    pad = (-nz) % align
    nzp = nz + pad
    nx, ny = 11, 9
    rng = np.random.default_rng(nz)
    mk = lambda: np.pad(rng.random((nx, ny, nz)).astype(dtype), ((0, 0), (0, 0), (0, pad)))
    ehs = [mk() for _ in range(6)]
    ces = [mk() for _ in range(3)]
    chs = [mk() for _ in range(3)]
    ref = O.ref_core(dtype, pad)
    a = [x.copy() for x in ehs]
    b = [x.copy() for x in ehs]
    for _ in range(3):
        ref.update_e(nx, ny, nzp, *[x.ctypes.data for x in a], *[c.ctypes.data for c in ces])
        ref.update_h(nx, ny, nzp, *[x.ctypes.data for x in a], *[c.ctypes.data for c in chs])
        O.update_e(b, ces, pad)
        O.update_h(b, chs, pad)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_core_linearity():
    """The update is linear in (E, H) for fixed CE: step(u+v) = step(u)+step(v) to rounding."""
    rng = np.random.default_rng(5)
    nx, ny, nz = 8, 9, 12
    mk = lambda: rng.standard_normal((nx, ny, nz))
    u = [mk() for _ in range(6)]
    v = [mk() for _ in range(6)]
    ces = [rng.random((nx, ny, nz)) for _ in range(3)]
    w = [x + y for x, y in zip(u, v)]
    for s in (u, v, w):
        O.update_e(s, ces, 0)
        O.update_h(s, None, 0)
    for x, y, z in zip(u, v, w):
        np.testing.assert_allclose(x + y, z, rtol=0, atol=1e-12)


# ------------------------------------------------------------------ CPML
def test_cpml_tables_closed_form():
    from fdtd_accelerate_amd.cpml import cpml_coefficients, sigma_max
    for npml, smax_expect in ((10, 3.4538776394910684), (12, 2.8782313662425572), (8, 4.317347049363836)):
        assert math.isclose(sigma_max(npml), smax_expect, rel_tol=1e-12)
        assert math.isclose(O.cpml_sigma_max(npml), smax_expect, rel_tol=1e-12)
        bE, aE, bH, aH = cpml_coefficients(npml)
        obE, oaE, obH, oaH = O.cpml_tables(npml, dtype=np.float64)
        for x, y in ((bE, obE), (aE, oaE), (bH, obH), (aH, oaH)):
            np.testing.assert_allclose(x, y, rtol=1e-15, atol=0)
        # sigma = 0 entries: E index 0; H indices 0 and npml+1 -> b = 1, a = 0
        assert bE[0] == 1 and aE[0] == 0
        assert bH[0] == 1 and aH[0] == 0 and bH[npml + 1] == 1 and aH[npml + 1] == 0
        # deepest cells: E front p=0 at depth npml-0.5, H front p=0 at depth npml
        smax = sigma_max(npml)
        assert math.isclose(bE[1], math.exp(-0.5 * smax * ((npml - 0.5) / npml) ** 4), rel_tol=1e-14)
        assert math.isclose(bH[1], math.exp(-0.5 * smax), rel_tol=1e-14)
        assert math.isclose(bH[2 * npml + 1], math.exp(-0.5 * smax), rel_tol=1e-14)
        # symmetric profile: front p <-> back npml-1-p
        np.testing.assert_array_equal(bE[1:npml + 1], bE[npml + 1:2 * npml + 1][::-1])
        np.testing.assert_array_equal(bH[1:npml + 1], bH[npml + 2:2 * npml + 2][::-1])
        assert np.all(aE[1:2 * npml + 1] < 0) and aE[2 * npml + 1] == 0 and np.all((bE > 0) & (bE <= 1))


def zero_walls(ehs):
    """Zero the cells the core never writes (tangential E on the high faces, tangential
    H on the low faces): the KEMP box's walls, which a physical field keeps at 0."""
    ex, ey, ez, hx, hy, hz = ehs
    ex[:, -1, :] = 0; ex[:, :, -1] = 0
    ey[-1, :, :] = 0; ey[:, :, -1] = 0
    ez[-1, :, :] = 0; ez[:, -1, :] = 0
    hx[:, 0, :] = 0; hx[:, :, 0] = 0
    hy[0, :, :] = 0; hy[:, :, 0] = 0
    hz[0, :, :] = 0; hz[:, 0, :] = 0


def _rand_sim(n, npml, seed, dtype=np.float64, nz=None):
    nz = nz or n
    rng = np.random.default_rng(seed)
    ces = [0.5 / rng.uniform(1, 4, (n, n, nz)) for _ in range(3)]
    s = O.OracleSim(n, n, nz, 0, npml, dtype, ces=ces)
    for a in s.ehs:
        a[:] = rng.standard_normal(a.shape)
    zero_walls(s.ehs)
    return s


def test_cpml_zero_a_is_noop():
    base = _rand_sim(16, 4, 1)
    pml = _rand_sim(16, 4, 1)
    pml.aE[:] = 0
    pml.aH[:] = 0
    pml.bE[:] = 1
    pml.bH[:] = 1
    base.npml = 0
    base.run(6)
    pml.run(6)
    for x, y in zip(base.ehs, pml.ehs):
        assert np.array_equal(x, y)
    for p1, p2 in list(pml.psiE.values()) + list(pml.psiH.values()):
        assert not p1.any() and not p2.any()


def test_cpml_cyclic_axis_symmetry():
    """x->y->z->x relabelling of a cubic CPML problem gives the same fields: all three
    face implementations (components, signs, psi indexing, table offsets) agree."""
    n, npml = 20, 4
    a = _rand_sim(n, npml, 7)
    ces = a.ces
    P = lambda x: np.ascontiguousarray(np.transpose(x, (2, 0, 1)))  # G[k,i,j] = F[i,j,k]
    b = O.OracleSim(n, n, n, 0, npml, np.float64, ces=[P(ces[2]), P(ces[0]), P(ces[1])])
    ex, ey, ez, hx, hy, hz = a.ehs
    for dst, src in zip(b.ehs, (P(ez), P(ex), P(ey), P(hz), P(hx), P(hy))):
        dst[:] = src
    a.run(12)
    b.run(12)
    ex, ey, ez, hx, hy, hz = a.ehs
    scale = max(np.abs(x).max() for x in a.ehs)
    for got, want in zip(b.ehs, (P(ez), P(ex), P(ey), P(hz), P(hx), P(hy))):
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * scale)
    # the CPML did act: psi on every face is non-zero
    for p1, p2 in list(a.psiE.values()) + list(a.psiH.values()):
        assert np.abs(p1).max() > 0 and np.abs(p2).max() > 0


def _cpml_1d(nx, npml, steps, src_i, tfunc, dtype=np.float32, pml=True):
    """Independent 1-D CPML (Ez at x=i+1/2, Hy at x=i; Hy(0) and Ez(nx-1) are walls) with
    sigma(x) = smax*(depth/npml)^4 from positions: front PML [0, npml], back PML
    [nx-1-npml, nx-1] (Roden & Gedney 2000 recursion, kappa=1, alpha=0, dt=0.5)."""
    smax = -5 * math.log(1e-6) / (2 * npml)

    def ba(depth):
        if not pml or depth <= 0:
            return 1.0, 0.0
        b = math.exp(-0.5 * smax * (depth / npml) ** 4)
        return b, b - 1.0

    bEx = np.ones(nx, dtype)
    aEx = np.zeros(nx, dtype)
    bHx = np.ones(nx, dtype)
    aHx = np.zeros(nx, dtype)
    back = nx - 1 - npml
    for i in range(nx):
        xe, xh = i + 0.5, float(i)
        de = npml - xe if xe < npml else (xe - back if xe > back else 0.0)
        dh = npml - xh if xh < npml else (xh - back if xh > back else 0.0)
        bEx[i], aEx[i] = ba(de) if i < nx - 1 else (1.0, 0.0)
        bHx[i], aHx[i] = ba(dh) if i > 0 else (1.0, 0.0)
    ez = np.zeros(nx, dtype)
    hy = np.zeros(nx, dtype)
    pe = np.zeros(nx, dtype)
    ph = np.zeros(nx, dtype)
    half = dtype(0.5)
    out_ez, out_hy = [], []
    for t in range(1, steps + 1):
        d = hy[1:] - hy[:-1]
        ez[:-1] = ez[:-1] + half * d
        pe[:-1] = bEx[:-1] * pe[:-1] + aEx[:-1] * d
        m = aEx[:-1] != 0
        ez[:-1][m] = ez[:-1][m] + half * pe[:-1][m]
        ez[src_i] += dtype(tfunc(t))
        d = ez[1:] - ez[:-1]
        hy[1:] = hy[1:] + half * d
        ph[1:] = bHx[1:] * ph[1:] + aHx[1:] * d
        m = aHx[1:] != 0
        hy[1:][m] = hy[1:][m] + half * ph[1:][m]
        out_ez.append(ez.copy())
        out_hy.append(hy.copy())
    return np.array(out_ez), np.array(out_hy)


def test_cpml_reduces_to_1d_plane_wave():
    """A plane source uniform in (y, z) with CPML on the x faces only: the centre line
    of the 3-D oracle equals the independent 1-D CPML while the y/z walls are outside
    the one-cell-per-step dependency cone."""
    nx, npml, steps = 40, 8, 70
    nyz = 2 * steps + 10
    tfunc = lambda t: O.pulse(t, 20.0, 6.0, 8.0)
    s = O.OracleSim(nx, nyz, nyz, 0, npml, np.float32,
                    cpml_faces=((True, True), (False, False), (False, False)))
    s.add_source("ez", (14, 0, 0), (14, nyz - 1, nyz - 1), tfunc)
    c = nyz // 2
    ez3, hy3 = [], []
    for _ in range(steps):
        s.step()
        ez3.append(s.field("ez")[:, c, c].copy())
        hy3.append(s.field("hy")[:, c, c].copy())
    ez1, hy1 = _cpml_1d(nx, npml, steps, 14, tfunc)
    scale = np.abs(ez1).max()
    assert scale > 0.5
    np.testing.assert_allclose(np.array(ez3), ez1, rtol=0, atol=2e-6 * scale)
    np.testing.assert_allclose(np.array(hy3), hy1, rtol=0, atol=2e-6 * scale)
    # the absorbed pulse entered both PMLs
    assert np.abs(np.array(ez3)[:, 2]).max() > 1e-3 and np.abs(np.array(ez3)[:, nx - 4]).max() > 1e-3


def test_cpml_1d_reflection_small():
    """Normal-incidence reflection of the 10-cell CPML (R=1e-6, m=4) for a pulse of the
    benchmark's band (period 20 steps), measured against a run in a domain long enough
    that nothing returns: about 1e-4 of the pulse (the bare wall reflects ~1)."""
    npml, steps = 10, 280
    tfunc = lambda t: O.pulse(t, 40.0, 10.0, 20.0)
    short, _ = _cpml_1d(80, npml, steps, 40, tfunc, np.float64)
    long_, _ = _cpml_1d(1260, npml, steps, 630, tfunc, np.float64)
    probe_s = short[:, 50]
    probe_l = long_[:, 640]
    refl = np.abs(probe_s - probe_l).max() / np.abs(probe_l).max()
    nopml, _ = _cpml_1d(80, npml, steps, 40, tfunc, np.float64, pml=False)
    refl_wall = np.abs(nopml[:, 50] - probe_l).max() / np.abs(probe_l).max()
    assert refl < 3e-4, refl
    assert refl_wall > 0.5  # without the CPML the wall reflects fully


def test_cpml_absorbs_3d_point_source():
    """Energy in a CPML-terminated box falls below 1e-6 of its peak once a zero-DC
    pulse (first difference of the benchmark-band pulse) has left; the same box
    without CPML (reflecting walls) keeps about a quarter of it."""
    n, npml, steps = 32, 8, 240
    tfunc = lambda t: O.pulse(t, 40.0, 10.0, 20.0) - O.pulse(t - 1, 40.0, 10.0, 20.0)

    def energy(s):
        return sum(float(np.sum(x.astype(np.float64) ** 2)) for x in s.ehs)

    a = O.OracleSim(n, n, n, 0, npml, np.float32)
    b = O.OracleSim(n, n, n, 0, 0, np.float32)
    for s in (a, b):
        s.add_source("ez", (16, 16, 16), (16, 16, 16), tfunc)
    ea, eb = [], []
    for _ in range(steps):
        a.step()
        b.step()
        ea.append(energy(a))
        eb.append(energy(b))
    assert max(ea) > 0.1
    assert ea[-1] < 1e-6 * max(ea), ea[-1] / max(ea)
    assert eb[-1] > 0.2 * max(eb)
    assert all(np.isfinite(ea))


# ------------------------------------------------------------------ sources
def test_pulse_closed_form():
    assert O.pulse(100.0) == 1.0
    assert math.isclose(O.pulse(110.0), -math.exp(-0.5 * (10 / 30) ** 2), rel_tol=1e-15)
    assert math.isclose(O.pulse(105.0), 0.0, abs_tol=1e-15)
    for t in (3.0, 50.0, 77.0):
        assert math.isclose(O.pulse(100 + t), O.pulse(100 - t), rel_tol=1e-14)
    from fdtd_accelerate_amd.sources import gaussian_pulse
    f = gaussian_pulse()
    for t in range(1, 300, 7):
        assert f(t) == O.pulse(float(t))


def test_source_tstep_rule():
    """K7/cpu/test/test_incident_direct.py:42-58: after one E+H step the box holds
    dtype(value) * dtype(tfunc(1)); the second E half uses tfunc(2)."""
    tf = lambda t: np.sin(0.03 * t)
    s2 = O.OracleSim(6, 7, 8, 0, 0, np.float32)
    s2.add_source("ex", (0, 6, 0), (0, 6, 7), tf, 0.37)  # ex at j = ny-1 is never updated by the core
    s2.step()
    want = np.float32(0.37) * np.float32(tf(1))
    assert np.all(s2.field("ex")[0, 6, :8] == want)
    s2.step()
    assert np.all(s2.field("ex")[0, 6, :8] == want + np.float32(0.37) * np.float32(tf(2)))
