"""GPU tests of the asynchronous snapshots (``SaveFields`` via ``YeeSim.set_savefields``,
K7/node/save_fields.py:10-48 and K7/main.py:239-317): the files written while the step
loop runs hold the fields of their step, byte-equal to the float32 oracle, on the
in-place and fused engines; several snapshot sets with their own ranges and tags."""
import os
import pickle

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
    from fdtd_accelerate_amd import _lib, snapshot
    _lib.load()
    snapshot.load()


def _oracle_states(n, npml, src, steps):
    """{step: [6 arrays without padding]} of the float32 oracle at every step."""
    nx, ny, nz = n
    o = O.OracleSim(nx, ny, nz, 0, npml, np.float32)
    o.add_source("ez", src, src, O.pulse)
    out = {}
    for t in range(1, steps + 1):
        o.step()
        out[t] = [a.copy() for a in o.ehs]
    return out


@pytest.mark.parametrize("engine", ["kemp", "fused"])
def test_savefields_match_oracle(tmp_path, engine):
    from fdtd_accelerate_amd.sim import YeeSim
    from fdtd_accelerate_amd.snapshot import read_dataset, dataset_info
    from fdtd_accelerate_amd.sources import gaussian_pulse
    n, npml, steps = (40, 36, 100), 6, 23
    src = (20, 18, 50)
    sim = YeeSim(*n, npml=npml, coeff_use="", engine=engine, max_tstep=steps)
    sim.add_source("ez", src, src, gaussian_pulse())
    d = str(tmp_path) + "/"
    whole = sim.set_savefields("ez", (0, 0, 0), (-1, -1, -1), (0, -1, 5), d)   # steps 5, 10, 15, 20
    plane = sim.set_savefields(["hy", "ex"], (7, 0, 0), (7, -1, -1), (3, 12, 2), d, tag=0,
                               nslots=1, nthreads=3)                           # steps 4 .. 12, even
    line = sim.set_savefields("hz", (0, 17, 50), (-1, 17, 50), (1, -1, 1), d, tag=5)  # every step
    assert (whole.tag, plane.tag, line.tag) == (0, 1, 5)
    sim.run(9)
    sim.run(steps - 9)
    sim.wait_savefields()
    ref = _oracle_states(n, npml, src, steps)
    want = {whole: [5, 10, 15, 20], plane: [4, 6, 8, 10, 12], line: list(range(1, steps + 1))}
    for sf, ts in want.items():
        files = sorted(f for f in os.listdir(d) if f.endswith(f"_tag{sf.tag}.h5"))
        assert files == [("%.2d_tag%d.h5" % (t, sf.tag)) for t in ts], (engine, sf.tag)
        for t in ts:
            p = d + "%.2d_tag%d.h5" % (t, sf.tag)
            for name in sf.str_fs:
                got = read_dataset(p, name)
                exp = ref[t][NAMES.index(name)][sf.slices]
                assert got.shape == sf.shape and dataset_info(p, name)[2] == 4
                assert got.tobytes() == np.ascontiguousarray(exp, np.float32).tobytes(), (engine, t, name)
    with open(d + "divide_info.pkl", "rb") as fh:   # the last set's, as the reference writes it
        info = pickle.load(fh)
    assert info["shape"] == (40,) and info["str_fs"] == ("hz",) and info["tgap"] == 1 and info["tmax"] == steps


def test_savefields_rejects():
    from fdtd_accelerate_amd.sim import YeeSim
    sim = YeeSim(16, 16, 16)
    with pytest.raises(ValueError, match="max_tstep"):
        sim.set_savefields("ez", (0, 0, 0), (-1, -1, -1), (0, -1, 5), "/tmp/")
    with pytest.raises(ValueError, match="tgap"):
        sim.set_savefields("ez", (0, 0, 0), (-1, -1, -1), (0, -1, 0), "/tmp/", max_tstep=10)
    with pytest.raises(ValueError):
        sim.set_savefields("ez", (0, 0, 0), (16, 0, 0), (0, -1, 1), "/tmp/", max_tstep=10)
    with pytest.raises(ValueError):
        sim.set_savefields("jx", (0, 0, 0), (1, 0, 0), (0, -1, 1), "/tmp/", max_tstep=10)


def test_geometry_h5_sizes_a_simulation(tmp_path):
    from fdtd_accelerate_amd.sim import YeeSim
    from fdtd_accelerate_amd.snapshot import write_geometry_h5
    p = str(tmp_path / "geo.h5")
    write_geometry_h5(p, 30, 28, 26, "e")
    sim = YeeSim.from_geometry_h5(p, max_tstep=100, npml=4)
    assert sim.fields.ns == [30, 28, 26] and sim.fields.ce_on and not sim.fields.ch_on
    assert sim.max_tstep == 100


def test_savefields_double_fields(tmp_path):
    """SaveFields on a float64 Fields (KEMP's precision_float='double', core updates only):
    float64 datasets equal to the float64 oracle at their steps."""
    from fdtd_accelerate_amd import Core, Fields
    from fdtd_accelerate_amd.snapshot import SaveFields, dataset_info, read_dataset
    nx, ny, nz = 10, 12, 30
    f = Fields(nx, ny, nz, coeff_use="e", precision_float="double")
    rng = np.random.default_rng(21)
    eh = [rng.standard_normal((nx, ny, nz)) for _ in range(6)]
    ce = [rng.random((nx, ny, nz)) for _ in range(3)]
    f.set_ehs(*eh)
    f.set_ces(*ce)
    Core(f)
    d = str(tmp_path) + "/"
    sf = SaveFields(f, ["ex", "hz"], (0, 0, 0), (nx - 1, ny - 1, nz - 1), d, 2, 100)
    pad = f.pad
    pz = lambda a, v=0: np.pad(a, ((0, 0), (0, 0), (0, pad)), constant_values=v)
    ref = [pz(x) for x in eh]
    rce = [pz(x, 0.5) for x in ce]
    want = {}
    for t in range(1, 7):
        f.update_e()
        f.update_h()
        O.update_e(ref, rce, pad)
        O.update_h(ref, None, pad)
        if t % 3 == 0:
            sf.save_fields(t)
            want[t] = (ref[0][:, :, :nz].copy(), ref[5][:, :, :nz].copy())
    sf.close()
    for t, (wx, wz) in want.items():
        p = d + "%.3d_tag2.h5" % t
        assert dataset_info(p, "ex")[1] == np.float64
        assert read_dataset(p, "ex").tobytes() == wx.tobytes()
        assert read_dataset(p, "hz").tobytes() == wz.tobytes()
