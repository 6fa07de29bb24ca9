"""CPU tests of the host side: argument checks with the reference's exceptions, index
helpers (the reference's doctest values), the C ABI library exports, media generator."""
import ctypes
import os
import re

import numpy as np
import pytest

from fdtd_accelerate_amd import util
from tests.conftest import ROOT


def test_check_type_and_value_errors():
    # K7/util/common.py:400-428 (TestFunctions)
    with pytest.raises(TypeError):
        util.check_type("", 2, float)
    with pytest.raises(TypeError):
        util.check_type("", 3.2, str)
    with pytest.raises(TypeError):
        util.check_type("", [1, 2], list)
    with pytest.raises(TypeError):
        util.check_type("", (1.2, 2.3), tuple, int)
    with pytest.raises(ValueError):
        util.check_value("", "a", ("b", "c"))
    util.check_type("a", (1.2, 2.3), tuple, float)
    util.check_value("a", "a", ("a", "b"))


def test_index_helpers_match_reference_doctests():
    assert util.convert_indices((10, 12, 14), (1, 2, 3)) == (1, 2, 3)
    assert util.convert_indices((10, 12, 14), (-1, -2, -3)) == (9, 10, 11)
    assert util.convert_indices((10, 12, 14), (0.1, 0.2, 0.3)) == (1, 2, 4)
    assert util.slices_two_points((0, 0, 0), (10, 11, 12)) == (slice(0, 11), slice(0, 12), slice(0, 13))
    assert util.slices_two_points((0, 0, 0), (10, 0, 12)) == (slice(0, 11), 0, slice(0, 13))
    assert util.slices_two_points((1, 2, 3), (1, 2, 3)) == (1, 2, 3)
    assert util.shape_two_points((0, 0, 0), (10, 11, 12)) == (11, 12, 13)
    assert util.shape_two_points((0, 0, 0), (10, 0, 12)) == (11, 13)
    assert util.shape_two_points((1, 2, 3), (1, 2, 3)) == (1,)
    assert util.shape_two_points((0, 0, 0), (10, 11, 12), 2) == (22, 12, 13)
    assert util.shape_two_points((0, 0, 0), (0, 0, 12), is_dummy=True) == (1, 1, 13)


def test_instance_priority_order():
    class I:
        def __init__(self, p):
            self.priority_type = p
    lst = []
    for p in ("incident", "core", "pml", "core", "incident"):
        util.append_instance(lst, I(p))
    assert [x.priority_type for x in lst] == ["core", "core", "pml", "incident", "incident"]


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "mifdtd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b(?:void|int|const char\s*\*)\s*\**\s*(\w+)\s*\(", txt)
    return sorted(set(names))


def test_library_exports_header_symbols():
    """libmifdtd.so loads without a GPU and exports every function of include/mifdtd.h."""
    from fdtd_accelerate_amd import _lib
    lib = _lib.load()
    names = _header_functions()
    assert len(names) >= 17
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_lib.EXPORTS)
    assert lib.mifdtd_version() >= 1


def test_argument_errors_without_gpu():
    """Rejected arguments come back as status codes before any device work."""
    from fdtd_accelerate_amd import _lib
    lib = _lib.load()
    rc = lib.mifdtd_update_e(4, 4, 6, 0, *([None] * 9), None)
    assert rc == -2 and "multiple of 4" in _lib.last_error()
    rc = lib.mifdtd_update_e(4, 4, 8, 0, *([None] * 9), None)
    assert rc == -3
    rc = lib.mifdtd_update_e(4, 4, 8, 7, *([16] * 6), None, None, None, None)
    assert rc == -1 and "pad" in _lib.last_error()
    rc = lib.mifdtd_update_h(4, 4, 8, 0, *([8] * 6), None, None, None, None)
    assert rc == -2 and "aligned" in _lib.last_error()
    d = _lib.CpmlHalf()
    d.npml = 4
    rc = lib.mifdtd_cpml_e(8, 8, 8, 0, d, *([16] * 6), None, None, None, None)
    assert rc == -1 and "too small for npml" in _lib.last_error()
    rc = lib.mifdtd_source(16, 4, 4, 8, 0, 4, 0, 0, 0, 0, 1.0, None, 1.0, 0, None)
    assert rc == -1 and "outside" in _lib.last_error()
    rc = lib.mifdtd_pbc(16, 32, 4, 4, 8, 0, 3, 1, None)
    assert rc == -1 and "axis" in _lib.last_error()
    rc = lib.mifdtd_pbc(16, 16, 4, 4, 8, 0, 0, 1, None)
    assert rc == -1 and "differ" in _lib.last_error()
    rc = lib.mifdtd_pbc(None, 16, 4, 4, 8, 0, 0, 1, None)
    assert rc == -3
    rc = lib.mifdtd_tfsf_line(16, 16, 16, 5, 1, 0.0, None)
    assert rc == -1 and "at least 6" in _lib.last_error()
    P = _lib.TfsfPatches()
    P.n = _lib.TFSF_MAX_PATCH + 1
    rc = lib.mifdtd_tfsf_faces(ctypes.byref(P), (ctypes.c_void_p * 6)(*[16] * 6), 16, 20, 8, 8, 8, None)
    assert rc == -1 and "patch count" in _lib.last_error()
    rc = lib.mifdtd_update_e_f64(4, 4, 7, 0, *([16] * 6), None, None, None, None)
    assert rc == -2 and "multiple of 2" in _lib.last_error()
    rc = lib.mifdtd_update_h_f64(4, 4, 6, 5, *([16] * 6), None, None, None, None)
    assert rc == -1 and "pad" in _lib.last_error()
    rc = lib.mifdtd_cpml_face_x_e(20, 20, 20, 0, 4, 2, *([16] * 10), None)
    assert rc == -1 and "backward" in _lib.last_error()
    with pytest.raises(RuntimeError):
        _lib.call("mifdtd_update_e", 4, 4, 6, 0, *([None] * 9), None)


def test_kemp_programs_export_reference_symbols():
    """Every KEMP core program library (precision x coeff_use x pad, plain and split form)
    loads without a GPU, exports update_e/update_h with the reference's arity (3 ints, or 5
    for the split form of K7/cpu/core_split.py) and names no other program."""
    from fdtd_accelerate_amd import core
    from fdtd_accelerate_amd.build import KEMP_COEFFS, KEMP_PRECISIONS
    n = 0
    for prec, pads in KEMP_PRECISIONS.items():
        for coeff in KEMP_COEFFS:
            for pad in pads:
                for split in (False, True):
                    lib = core.program(prec, coeff, pad, split=split)
                    ni = 5 if split else 3
                    assert len(lib.update_e.argtypes) == ni + (9 if "e" in coeff else 6)
                    assert len(lib.update_h.argtypes) == ni + (9 if "h" in coeff else 6)
                    n += 1
    assert n == 48
    with pytest.raises(ValueError):
        core.program("double", "e", 2)
    with pytest.raises(ValueError):
        core.program("half", "e", 0)
    # the main library exports no reference-named core symbol (the programs do)
    from fdtd_accelerate_amd import _lib
    import ctypes as C
    main = C.CDLL(_lib.LIB_PATH)
    for name in ("update_e", "update_h", "mifdtd_set_pad", "update_x_e"):
        assert not hasattr(main, name), name


def test_kemp_program_rejects_on_cpu_through_last_error():
    """A rejected call through a program records its reason (the void ABI's only error
    channel): nz_pitch 6 is not a multiple of 4 for a single program."""
    from fdtd_accelerate_amd import _lib, core
    prog = core.program("single", "", 2)
    prog.update_e(4, 4, 6, *([None] * 6))
    assert "multiple of 4" in _lib.last_error()
    prog = core.program("double", "eh", 1)
    prog.update_h(4, 4, 7, *([None] * 9))
    assert "multiple of 2" in _lib.last_error()
    # split form: idx0 must start a vector and the range must lie in the grid (checked
    # before any launch; 64 stands for an aligned device pointer)
    prog = core.program("single", "", 0, split=True)
    prog.update_e(4, 4, 8, 2, 100, *([64] * 6))
    assert "does not start a 4-element vector" in _lib.last_error()
    prog.update_h(4, 4, 8, 0, 129, *([64] * 6))
    assert "outside the grid" in _lib.last_error()
    prog = core.program("double", "h", 1, split=True)
    prog.update_h(4, 4, 8, 1, 100, *([64] * 9))
    assert "does not start a 2-element vector" in _lib.last_error()


def test_fields_requires_gpu():
    import torch
    from fdtd_accelerate_amd import Fields
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(RuntimeError):
        Fields(8, 8, 8)
    with pytest.raises(RuntimeError):
        Fields(8, 8, 8, precision_float="double")
    with pytest.raises(ValueError):
        Fields(8, 8, 8, precision_float="quadruple")
    with pytest.raises(TypeError):
        Fields(8.0, 8, 8)


def test_random_spheres_deterministic_and_disjoint():
    from fdtd_accelerate_amd.geometry import random_spheres
    a = random_spheres(3, 30, (3, 8), (1, 12), (10, 10, 10), (90, 90, 90))
    b = random_spheres(3, 30, (3, 8), (1, 12), (10, 10, 10), (90, 90, 90))
    assert [(s.center, s.radius, s.eps_r) for s in a] == [(s.center, s.radius, s.eps_r) for s in b]
    for i, s in enumerate(a):
        c = np.array(s.center)
        assert np.all(c - s.radius >= 10) and np.all(c + s.radius <= 90)
        assert 1 <= s.eps_r <= 12 and 3 <= s.radius <= 8
        for t in a[:i]:
            assert np.linalg.norm(c - np.array(t.center)) >= s.radius + t.radius


def test_paint_ce_at_yee_points():
    """CE = 0.5/eps at each component's staggered point: a sphere centred on an Ex
    point includes it; a point just outside the radius keeps the background."""
    import torch
    from fdtd_accelerate_amd.geometry import Sphere, paint_ce
    ces = [torch.zeros((20, 20, 20)) for _ in range(3)]
    s = Sphere((10.0, 10.5, 10.5), 3.2, 4.0)
    paint_ce(ces, [s], (20, 20, 20))
    assert ces[0][10, 10, 10] == np.float32(0.125)       # Ex(10,10,10) at (10, 10.5, 10.5): centre
    assert ces[0][13, 10, 10] == np.float32(0.125)       # distance 3 < 3.2
    assert ces[0][14, 10, 10] == np.float32(0.5)         # distance 4
    assert ces[1][10, 10, 10] == np.float32(0.125)       # Ey at (10.5, 10, 10.5): distance ~0.71
    assert ces[2][11, 12, 10] == np.float32(0.125)       # Ez at (11.5, 12.5, 10): distance ~2.55
    assert ces[2][13, 12, 10] == np.float32(0.5)         # Ez at (13.5, 12.5, 10): distance ~4.06
    n_in = int((ces[0] == np.float32(0.125)).sum())
    assert abs(n_in - 4 / 3 * np.pi * 3.2 ** 3) < 0.25 * 4 / 3 * np.pi * 3.2 ** 3


def test_adjoint_design_box_off_the_cpml():
    """The adjoint gradient has no term for the CPML correction of E, so a design box
    touching a CPML slab is rejected by the C entry point and by DesignProblem."""
    from fdtd_accelerate_amd import _lib
    lib = _lib.load()
    lam = (ctypes.c_void_p * 6)(*([16] * 6))
    ce = (ctypes.c_void_p * 3)(None, None, None)
    for box, ok in (((10, 10, 10, 19, 19, 19), True), ((9, 10, 10, 19, 19, 19), False),
                    ((10, 10, 10, 19, 19, 20), False), ((10, 19, 10, 19, 19, 19), False)):
        b = _lib.DesignBox(*box)
        # 40^3, npml 10: E slab planes are [0, 10) and [29, 39) on each axis
        rc = lib.mifdtd_adjoint_step(40, 40, 40, 0, lam, ce, 10, 16, 16, 16, 16, 16, 16, None, 0,
                                     ctypes.byref(b), None if ok else 16, 16, None)
        if ok:
            assert rc == -3 and "curl" in _lib.last_error()  # passes the box check, then needs curl
        else:
            assert rc == -1 and "overlaps a CPML slab" in _lib.last_error(), box
    from fdtd_accelerate_amd.adjoint import DesignProblem
    with pytest.raises((ValueError, RuntimeError)) as e:
        DesignProblem(n=(40, 40, 40), npml=10, box=(8, 12, 12, 10, 10, 10))
    assert e.type is ValueError or "GPU" in str(e.value)


def test_kufdtd_shaped_psi_is_rejected():
    """A face psi allocated with kuFDTD's shapes (KU/cpu/cpml.py:35-85) is refused before
    any device call, since the face kernels index this engine's shapes."""
    import torch
    from fdtd_accelerate_amd.cpml import check_face_psi, psi_shape

    class F:  # a Fields stand-in: only the pitch is read
        ns_pitch = (30, 32, 36)
    nx, ny, nz, npml = 30, 32, 36, 6
    kufdtd = [(npml, ny - 1, nz), (nx - 1, npml, nz), (nx - 1, ny - 1, npml + 1)]
    for axis in range(3):
        bad = torch.zeros(kufdtd[axis])
        with pytest.raises(ValueError, match="must be a contiguous float32 tensor"):
            check_face_psi(F, axis, npml, bad, bad)
        good = torch.zeros(psi_shape(axis, nx, ny, nz, npml))
        check_face_psi(F, axis, npml, good, good)
        with pytest.raises(ValueError):
            check_face_psi(F, axis, npml, good, good.double())


def test_static_batch_rejects_beyond_its_limit_without_gpu():
    """The allocation-free batched run refuses more instances than its static array holds,
    before any device work."""
    from fdtd_accelerate_amd import _lib
    descs = (_lib.InstanceDesc * 129)()
    rc = _lib.load().mifdtd_batch_run_static(descs, 129, None, None, None, None, 1, None, None, 0, None)
    assert rc == -4 and "at most 128" in _lib.last_error()


def test_batch_ce_index_arguments_rejected_without_gpu():
    """CE material indices on some instances only, an incomplete index set or a
    misaligned index array: refused before any device work (field pointers are fake,
    16-byte aligned host values)."""
    from fdtd_accelerate_amd import _lib
    lib = _lib.load()
    descs = (_lib.InstanceDesc * 2)()
    for d in descs:
        d.nx, d.ny, d.nz_pitch, d.pad = 8, 8, 8, 0
        for c in range(6):
            d.e[c] = 4096 * (c + 1)
    out = ctypes.c_void_p()
    descs[0].ce_mat[0] = descs[0].ce_mat[1] = descs[0].ce_mat[2] = 65536
    descs[0].ce_lut = 65536
    rc = lib.mifdtd_batch_create(descs, 2, None, None, None, None, ctypes.byref(out))
    assert rc == -3 and "instance 1" in _lib.last_error() and "every instance" in _lib.last_error()
    descs[1].ce_mat[0] = descs[1].ce_mat[1] = 65536
    rc = lib.mifdtd_batch_create(descs, 2, None, None, None, None, ctypes.byref(out))
    assert rc == -3 and "all three" in _lib.last_error()
    descs[1].ce_mat[2], descs[1].ce_lut = 65538, 65536
    rc = lib.mifdtd_batch_create(descs, 2, None, None, None, None, ctypes.byref(out))
    assert rc == -2 and "4-byte aligned" in _lib.last_error()
    assert not out.value


def test_z_line_phase_choice():
    """The fused engine's line phase: a multiple of 4 (16-byte quads) putting the sweep's
    z range [kb, ke) on line boundaries (k + phase a multiple of 32) inside the CPML-free
    cells [npml+1, nz-1-npml), with the fewest cells left between (checked against every
    phase, and against a brute-force scan of line boundaries); 0 when rows of nz_pitch
    floats do not share one phase."""
    from fdtd_accelerate_amd.fused import z_line_phase
    assert z_line_phase(1024, 1024, 12) == 16
    assert z_line_phase(1000, 1008, 4) == 0 and z_line_phase(40, 48, 4) == 0

    def box(nz, p, ph):  # brute force: the first and last line boundaries in the CPML-free range
        lines = [k for k in range(nz + 1) if (k + ph) % 32 == 0 and p + 1 <= k <= nz - 1 - p]
        return (lines[0], lines[-1]) if len(lines) >= 2 else None

    for nz, nzp, p in [(1024, 1024, 12), (160, 160, 6), (416, 416, 6), (512, 512, 10), (128, 128, 10),
                       (57, 64, 5), (300, 320, 5), (96, 96, 8)]:
        ph = z_line_phase(nz, nzp, p)
        assert ph % 4 == 0 and 0 <= ph < 32
        waste = {q: (nz - 2 - 2 * p) - (b[1] - b[0]) for q in range(0, 32, 4) if (b := box(nz, p, q))}
        if not waste:
            assert ph == 0
            continue
        assert ph in waste and waste[ph] == min(waste.values()), (nz, p, ph, waste)


def test_fused_shell_boxes_cover_the_grid(tmp_path):
    """The fused step's CPML shell decomposition (csrc/shell_boxes.hpp, as the kernels use
    it, built here for the host): for 10 grids, CPML widths 0-12, every line phase and
    with or without the step-pair fringes, the shell boxes decoded thread by thread cover
    each quad outside the interior box exactly once, none inside it, and never leave the
    grid (tests/shell_cover.cpp)."""
    import subprocess
    src = os.path.join(os.path.dirname(__file__), "shell_cover.cpp")
    exe = str(tmp_path / "shell_cover")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-std=c++17", "-O1", "-I", os.path.join(root, "include"),
                    src, "-o", exe], check=True, capture_output=True, timeout=600)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and "shell cover ok" in out.stdout, out.stdout + out.stderr


def test_ctypes_mirrors_match_header(tmp_path):
    """Every ctypes mirror in _lib has the size and field offsets of its C struct in
    include/mifdtd.h (a C probe compiled with gcc against the header prints them)."""
    import ctypes
    import shutil
    import subprocess
    from fdtd_accelerate_amd import _lib
    if shutil.which("gcc") is None:
        pytest.skip("no gcc")
    pairs = [(_lib.CpmlHalf, "mifdtd_cpml_half"), (_lib.TfsfPatches, "mifdtd_tfsf_patches"),
             (_lib.FusedDesc, "mifdtd_fused_desc"), (_lib.InstanceDesc, "mifdtd_instance"),
             (_lib.DesignBox, "mifdtd_design_box")]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mifdtd.h"', "int main(void) {"]
    for cls, cname in pairs:
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "probe"
    inc = os.path.join(ROOT, "include")
    subprocess.run(["gcc", "-std=c11", "-I", inc, "-I", "/opt/rocm/include", str(src), "-o", str(exe)], check=True,
                   capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cls, cname in pairs:
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)


@pytest.mark.parametrize("argv, env", [(["--gpus", "2"], {}), ([], {"WORLD_SIZE": "8"})])
def test_bench_refuses_more_than_one_gpu(argv, env):
    """bench.py measures one GPU: asked for N > 1 (by flag or by a launcher's WORLD_SIZE)
    it exits with status 2 and a message, before touching torch or a device."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *argv], capture_output=True, text=True,
                       env={**os.environ, **env}, timeout=120)
    assert r.returncode == 2 and r.stdout == ""
    assert "one GPU only" in r.stderr
