"""GPU parity of the 4th-order symplectic path against the oracle (bit-exact FP32).

Stage kernels (mifdtd_s4_update_h/e and the reference-named s4_update_h/e) on random
fields with random coefficient arrays and with the constant-coefficient form; the full
step (5 H stages, 4 E stages, CPML after each stage, sources after each E stage at
tstep + s/5, probes) against OracleS4Sim."""
import ctypes

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


def _dev(a):
    """(nx, ny, nz) Fortran-order host array -> device tensor (nz, ny, nx)."""
    return torch.tensor(np.ascontiguousarray(np.asarray(a).transpose(2, 1, 0)), device="cuda")


def _host(t):
    return np.asfortranarray(t.cpu().numpy().transpose(2, 1, 0))


@pytest.mark.parametrize("shape", [(8, 9, 10), (16, 7, 6), (32, 20, 24), (12, 33, 17)])
def test_s4_stages_match_oracle(shape):
    from fdtd_accelerate_amd import _lib
    rng = np.random.default_rng(sum(shape))
    a = [np.asfortranarray(rng.standard_normal(shape).astype(np.float32)) for _ in range(6)]
    c = [np.asfortranarray(rng.uniform(0.01, 0.04, shape).astype(np.float32)) for _ in range(6)]
    d = [_dev(x) for x in a]
    dc = [_dev(x) for x in c]
    nx, ny, nz = shape
    st = _lib.stream_handle(torch.cuda.current_stream())
    P = [t.data_ptr() for t in d]
    for s in range(4):
        O.s4_update_h(a, c[3:], O.S4_C[s])
        _lib.call("mifdtd_s4_update_h", nx, ny, nz, float(np.float32(O.S4_C[s])), *P,
                  *[t.data_ptr() for t in dc[3:]], 0.0, st)
        O.s4_update_e(a, c[:3], O.S4_D[s])
        _lib.call("mifdtd_s4_update_e", nx, ny, nz, float(np.float32(O.S4_D[s])), *P,
                  *[t.data_ptr() for t in dc[:3]], 0.0, st)
    torch.cuda.synchronize()
    for name, x, t in zip("ex ey ez hx hy hz".split(), a, d):
        assert x.tobytes() == _host(t).tobytes(), name


def test_s4_reference_named_and_constant_forms():
    """s4_update_h/e (reference argument order, arrays) and the NULL-array form with a
    constant coefficient both equal the oracle fed constant arrays."""
    from fdtd_accelerate_amd import _lib
    shape = (20, 14, 12)
    nx, ny, nz = shape
    rng = np.random.default_rng(3)
    init = [np.asfortranarray(rng.standard_normal(shape).astype(np.float32)) for _ in range(6)]
    c0 = np.float32(0.6 / 24)
    const = [np.full(shape, c0, np.float32, order="F") for _ in range(3)]
    a = [x.copy(order="F") for x in init]
    d1 = [_dev(x) for x in init]
    d2 = [_dev(x) for x in init]
    dc = [_dev(x) for x in const]
    st = _lib.stream_handle(torch.cuda.current_stream())
    lib = _lib.load()
    for s in range(2):
        O.s4_update_h(a, const, O.S4_C[s])
        O.s4_update_e(a, const, O.S4_D[s])
        lib.s4_update_h(nx, ny, nz, float(np.float32(O.S4_C[s])), *[t.data_ptr() for t in d1], *[t.data_ptr() for t in dc])
        lib.s4_update_e(nx, ny, nz, float(np.float32(O.S4_D[s])), *[t.data_ptr() for t in d1], *[t.data_ptr() for t in dc])
        _lib.call("mifdtd_s4_update_h", nx, ny, nz, float(np.float32(O.S4_C[s])), *[t.data_ptr() for t in d2],
                  None, None, None, float(c0), st)
        _lib.call("mifdtd_s4_update_e", nx, ny, nz, float(np.float32(O.S4_D[s])), *[t.data_ptr() for t in d2],
                  None, None, None, float(c0), st)
    torch.cuda.synchronize()
    for x, t1, t2 in zip(a, d1, d2):
        assert x.tobytes() == _host(t1).tobytes()
        assert x.tobytes() == _host(t2).tobytes()


@pytest.mark.parametrize("shape,npml,steps", [
    ((24, 22, 20), 4, 6),
    ((32, 28, 36), 6, 5),
    ((40, 17, 23), 5, 4),
])
def test_s4_step_matches_oracle(shape, npml, steps):
    """Full steps with CPML, a dielectric CE map, a point and a plane source and probes."""
    from fdtd_accelerate_amd.s4 import S4Sim
    from fdtd_accelerate_amd.sources import gaussian_pulse
    nx, ny, nz = shape
    S = 0.65
    rng = np.random.default_rng(nx)
    eps = rng.uniform(1, 6, shape).astype(np.float32)
    ce = (np.float32(S / 24) / eps).astype(np.float32)
    init = [np.asfortranarray(rng.standard_normal(shape).astype(np.float32)) for _ in range(6)]
    tf = gaussian_pulse(4.0, 3.0, 8.0)
    sv = rng.standard_normal((1, ny - 6, nz - 6)).astype(np.float32)
    o = O.OracleS4Sim(nx, ny, nz, S, npml=npml, ces=[ce, ce, ce])
    g = S4Sim(nx, ny, nz, S=S, npml=npml)
    g.set_ce(ce, ce, ce)
    for k, name in enumerate("ex ey ez hx hy hz".split()):
        o.ehs[k][:] = init[k]
        g.set_field(name, init[k])
    for sim in (o, g):
        sim.add_source("ez", (nx // 2, ny // 2, nz // 2), (nx // 2, ny // 2, nz // 2), tf)
        sim.add_source("ey", (npml + 2, 3, 3), (npml + 2, ny - 4, nz - 4), tf, sv)
        sim.add_probe("ez", (nx // 2 + 2, ny // 2, nz // 2))
        sim.add_probe("hy", (npml, 5, 6))
    series = g.run(steps)
    o.run(steps)
    for k, name in enumerate("ex ey ez hx hy hz".split()):
        assert o.ehs[k].tobytes() == np.asfortranarray(g.field(name).cpu().numpy()).tobytes(), name
    assert series.tobytes() == np.array(o.series, np.float32).tobytes()
    for key in o.psiE:
        for c in range(2):
            for od, gd in ((o.psiE, g.psiE), (o.psiH, g.psiH)):
                ref = od[key][c]
                got = np.asfortranarray(gd[key][c].cpu().numpy().transpose(2, 1, 0))
                assert ref.tobytes() == got.tobytes(), (key, c)


def test_s4_spheres_64cube_vs_oracle():
    """The 4th-order sphere workload at 64^3: 0.9 S_max, 10-cell CPML, sphere dielectric, Ez point source with
    the Gaussian-modulated pulse, 150 steps: probe series and fields equal the float32 oracle."""
    from fdtd_accelerate_amd import geometry
    from fdtd_accelerate_amd.s4 import S4Sim, courant_limit
    n, npml, steps = 64, 10, 150
    S = 0.9 * courant_limit()
    spheres = geometry.random_spheres(77, 6, (3.0, 8.0), (1.0, 12.0), (npml,) * 3, (n - npml,) * 3)
    eps = geometry.eps_map_numpy((n, n, n), spheres)
    o = O.OracleS4Sim(n, n, n, S, npml=npml, ces=[(np.float32(S / 24) / eps).astype(np.float32)] * 3)
    g = S4Sim(n, n, n, S=S, npml=npml)
    g.set_eps(eps)
    for sim in (o, g):
        sim.add_source("ez", (32, 32, 32), (32, 32, 32), O.pulse)
        sim.add_probe("ez", (44, 32, 32))
    series = g.run(steps)
    o.run(steps)
    assert series.tobytes() == np.array(o.series, np.float32).tobytes()
    for k, name in enumerate("ex ey ez hx hy hz".split()):
        assert o.ehs[k].tobytes() == np.asfortranarray(g.field(name).cpu().numpy()).tobytes(), name
    # sanity: the pulse arrived at the probe
    assert np.abs(series).max() > 1e-4


def test_s4_rejects_bad_shapes():
    from fdtd_accelerate_amd import _lib
    from fdtd_accelerate_amd.s4 import S4Sim
    with pytest.raises(ValueError):
        S4Sim(18, 16, 16)
    t = torch.zeros(16 * 16 * 18, device="cuda")
    with pytest.raises(RuntimeError):
        _lib.call("mifdtd_s4_update_e", 18, 16, 16, 0.5, *([t.data_ptr()] * 6), None, None, None, 0.01,
                  _lib.stream_handle(torch.cuda.current_stream()))


def test_s4_ce_table_equals_float_arrays():
    """E stages reading CE as uint8 material indices + table (set_eps on a sphere medium)
    equal the float-array E stages bit for bit over a full run with CPML and a source."""
    from fdtd_accelerate_amd import geometry
    from fdtd_accelerate_amd.s4 import S4Sim
    from fdtd_accelerate_amd.sources import gaussian_pulse
    n, npml = 52, 6
    spheres = geometry.random_spheres(9, 12, (2.0, 5.0), (1.0, 12.0), (npml,) * 3, (n - npml,) * 3)
    eps = geometry.eps_map_numpy((n, n, n), spheres)
    sims = [S4Sim(n, n, n, npml=npml) for _ in range(2)]
    for s in sims:
        s.set_eps(eps)
        s.add_source("ez", (26, 26, 26), (26, 26, 26), gaussian_pulse(10.0, 4.0, 8.0))
        s.add_probe("ez", (30, 26, 26))
    assert sims[0].ce_mat is not None
    sims[1].ce_mat = sims[1].ce_lut = None  # float arrays
    out = [s.run(40) for s in sims]
    assert out[0].tobytes() == out[1].tobytes()
    for name in ("ex", "ey", "ez", "hx", "hy", "hz"):
        assert sims[0].field(name).cpu().numpy().tobytes() == sims[1].field(name).cpu().numpy().tobytes(), name
