"""CPU tests of the HDF5 snapshot and geometry files (libmifdtd_h5.so, no GPU).

Pins:
* the reader on the reference's own geometry file (tests/golden/kemp_geo.h5, a copy of
  KEMP/v0.7_pml/fdtd3d/test/geo.h5, which K7/test/create_sample_h5.py:3-8 made with
  nx=150, ny=140, nz=64, coeff_use='');
* the geometry writer against that file: the HDF5 library's own h5dump prints the same
  attributes, types and values for both;
* the snapshot writer through the HDF5 library's deflate filter: our chunks are
  compressed by zlib on several threads, and reading them back through the library's
  filter pipeline (and h5dump) gives the input bit for bit, at ragged chunk edges too;
* the file names of K7/node/save_fields.py:30-31 on hand-derived examples.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from fdtd_accelerate_amd import snapshot as S

GOLD = os.path.join(os.path.dirname(__file__), "golden", "kemp_geo.h5")
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)
need_h5dump = pytest.mark.skipif(H5DUMP is None, reason="h5dump (HDF5 tools) not installed")


def _dump(path, *args):
    r = subprocess.run([H5DUMP, *args, path], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return r.stdout.split("\n", 1)[1]  # drop the line naming the file


def test_library_exports_header_symbols():
    """libmifdtd_h5.so exports every function include/mifdtd_h5.h declares, and the
    ctypes table binds exactly those."""
    txt = open(os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "mifdtd_h5.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"\b(?:int|const char\s*\*)\s*\**\s*(\w+)\s*\(", txt))
    lib = S.load()
    assert len(names) == 6 and names == set(S.EXPORTS)
    for name in names:
        assert hasattr(lib, name), name


def test_reads_reference_geometry_file():
    assert S.read_geometry_h5(GOLD) == ("", 150, 140, 64)


@need_h5dump
def test_geometry_writer_matches_reference_file(tmp_path):
    p = str(tmp_path / "geo.h5")
    S.write_geometry_h5(p, 150, 140, 64, "")
    assert _dump(p) == _dump(GOLD)
    for cu in ("e", "h", "eh"):
        S.write_geometry_h5(p, 7, 8, 9, cu)
        assert S.read_geometry_h5(p) == (cu, 7, 8, 9)
        assert f'STRSIZE {len(cu)};' in _dump(p) and f'"{cu}"' in _dump(p)


def test_geometry_errors(tmp_path):
    p = str(tmp_path / "not_h5.h5")
    with open(p, "w") as fh:
        fh.write("plain text")
    with pytest.raises(RuntimeError, match="cannot open"):
        S.read_geometry_h5(p)
    q = str(tmp_path / "snap.h5")
    S.write_fields(q, {"ez": np.zeros((2, 3), np.float32)})
    with pytest.raises(RuntimeError, match="attribute 'nx' is missing"):
        S.read_geometry_h5(q)
    with pytest.raises(ValueError):
        S.write_geometry_h5(p, 1, 2, 3, "x")


# (shape, dtype): one chunk; two chunks, the second ragged; rank 2 in four chunks, the
# last ragged; rank 1 in two ragged chunks; a single value
_CASES = [((5, 4, 3), np.float32), ((71, 64, 65), np.float32), ((301, 1500), np.float64),
          ((300001,), np.float32), ((1,), np.float64)]


@pytest.mark.parametrize("shape,dtype", _CASES)
@pytest.mark.parametrize("nthreads", [1, 7])
def test_fields_roundtrip(tmp_path, shape, dtype, nthreads):
    rng = np.random.default_rng(sum(shape) + nthreads)
    data = {n: rng.standard_normal(shape).astype(dtype) for n in ("ex", "hz")}
    data["hz"][..., :1] = -0.0  # signed zeros survive
    p = str(tmp_path / "snap.h5")
    S.write_fields(p, data, gzip_level=4, nthreads=nthreads)
    for n, a in data.items():
        shp, dt, lvl = S.dataset_info(p, n)
        assert shp == shape and dt == dtype and lvl == 4
        got = S.read_dataset(p, n)
        assert got.dtype == dtype and got.tobytes() == a.tobytes(), n


@need_h5dump
def test_h5dump_sees_gzip_datasets(tmp_path):
    rng = np.random.default_rng(3)
    a = rng.standard_normal((4, 3, 5)).astype(np.float32)
    p = str(tmp_path / "0100_tag0.h5")
    S.write_fields(p, {"ez": a}, gzip_level=6)
    hdr = _dump(p, "-p", "-H")
    assert "DATASET \"ez\"" in hdr and "H5T_IEEE_F32LE" in hdr
    assert "SIMPLE { ( 4, 3, 5 ) / ( 4, 3, 5 ) }" in hdr
    assert re.search(r"COMPRESSION DEFLATE \{ LEVEL 6 \}", hdr)
    body = _dump(p, "-d", "/ez", "-m", "%.9g", "-y", "-w", "0")
    vals = np.array([float(v) for v in re.findall(r"[-+]?[0-9.]+(?:e[-+]?[0-9]+)?",
                                                  body.split("DATA {", 1)[1].split("}", 1)[0])], np.float32)
    assert vals.tobytes() == a.ravel().tobytes()


def test_write_rejects(tmp_path):
    p = str(tmp_path / "x.h5")
    with pytest.raises(TypeError):
        S.write_fields(p, {"ez": np.zeros(4, np.int32)})
    with pytest.raises(ValueError):
        S.write_fields(p, {"ez": np.zeros(4, np.float32), "ex": np.zeros(5, np.float32)})
    with pytest.raises(RuntimeError, match="rank"):
        S.write_fields(p, {"ez": np.zeros((1, 1, 1, 2), np.float32)})
    with pytest.raises(RuntimeError, match="gzip level"):
        S.write_fields(p, {"ez": np.zeros(4, np.float32)}, gzip_level=10)
    with pytest.raises(RuntimeError, match="cannot create"):
        S.write_fields(str(tmp_path / "no_dir" / "x.h5"), {"ez": np.zeros(4, np.float32)})
    S.write_fields(p, {"ez": np.zeros(4, np.float32)})
    with pytest.raises(RuntimeError, match="no dataset 'ex'"):
        S.read_dataset(p, "ex")
    buf = np.zeros(2)
    assert S.load().mifdtd_h5_read_dataset(p.encode(), b"ez", buf.ctypes.data, buf.nbytes // 2, 1) != 0
    assert "not 8 bytes" in S.load().mifdtd_h5_last_error().decode()
    assert S.read_dataset(p, "ez", np.float64).dtype == np.float64  # converted on read


def test_file_names():
    """K7/node/save_fields.py:30-31: one digit more than the exponent of max_tstep."""
    assert S.fpath_form("out/", 0, 1000) % 100 == "out/0100_tag0.h5"
    assert S.fpath_form("out/", 3, 999) % 7 == "out/007_tag3.h5"
    assert S.fpath_form("", 1, 9) % 9 == "9_tag1.h5"
    assert S.fpath_form("d/", 2, 10) % 5 == "d/05_tag2.h5"
    assert S.fpath_form("d/", 0, 123456) % 42 == "d/000042_tag0.h5"


def test_divide_info(tmp_path):
    import pickle
    d = str(tmp_path) + "/"
    S.write_divide_info(d, (11, 13), (0, 2, 5), (10, 2, 17), 40, ("ez",), 1000, 10)
    with open(d + "divide_info.pkl", "rb") as fh:
        info = pickle.load(fh)  # written by this package
    assert info == {"shape": (11, 13), "pt0": (0, 2, 5), "pt1": (10, 2, 17), "anx_list": [0, 39],
                    "str_fs": ("ez",), "tmax": 1000, "tgap": 10}
