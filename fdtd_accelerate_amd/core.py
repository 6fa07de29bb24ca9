"""KEMP ``Core`` (``K7/cpu/core.py:8-123``, ``K7/gpu/core.py``): the Yee curl updates of a
``Fields`` through the HIP kernels of ``libmifdtd.so``; ``CoreSplit``, the split-form
core of ``K7/cpu/core_split.py`` (update_e / update_h of a part of the grid, for
computation/communication overlap); and ``program()``, the KEMP core program library of
one (precision_float, coeff_use, pad)."""
from __future__ import annotations

import ctypes
import os

from . import _lib, util
from .build import KEMP_COEFFS, KEMP_PRECISIONS, kemp_program_path
from .fields import Fields

_programs = {}


def program(precision_float="single", coeff_use="e", pad=0, split=False):
    """Load the KEMP core program of (precision_float, coeff_use, pad): the library that
    ``K7/cpu/core.py:52-85`` would generate and load for those options, with its two
    symbols bound as that file binds them (``:89-105``): ``update_e`` / ``update_h``
    take 3 ``c_int`` (nx, ny, nz_pitch) and 6 arrays, plus the 3 CE (CH) arrays when
    ``coeff_use`` holds 'e' ('h'); ``restype`` None.  Arrays are device pointers
    (``ctypes.c_void_p``, e.g. ``tensor.data_ptr()``) instead of the reference's host
    ``ndpointer``.  The pad is fixed per program, as the reference's masks are.
    ``split``: the program of ``K7/cpu/core_split.py:89-98``, whose two symbols take
    ``idx0, nmax`` (``c_int``) after nz_pitch and update that flat element range."""
    util.check_value("precision_float", precision_float, tuple(KEMP_PRECISIONS))
    util.check_value("coeff_use", coeff_use, KEMP_COEFFS)
    util.check_value("pad", pad, KEMP_PRECISIONS[precision_float])
    key = (precision_float, coeff_use, pad, bool(split))
    if key not in _programs:
        _lib.load()  # the program forwards into libmifdtd.so: load it first, by path
        path = kemp_program_path(*key)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built: run `python -m fdtd_accelerate_amd.build`")
        lib = ctypes.CDLL(path)
        arr = ctypes.c_void_p
        ints = [ctypes.c_int] * (5 if split else 3)
        lib.update_e.argtypes = ints + [arr] * (9 if "e" in coeff_use else 6)
        lib.update_e.restype = None
        lib.update_h.argtypes = ints + [arr] * (9 if "h" in coeff_use else 6)
        lib.update_h.restype = None
        lib.mifdtd_kemp_program_id.restype = ctypes.c_int
        want = (int(bool(split)) << 16) | (int(precision_float == "double") << 12) | (int("e" in coeff_use) << 8) | \
            (int("h" in coeff_use) << 4) | pad
        if lib.mifdtd_kemp_program_id() != want:
            raise RuntimeError(f"{path} is not the program {key}")
        _programs[key] = lib
    return _programs[key]


class Core:
    priority_type = "core"

    def __init__(self, fields):
        util.check_type("fields", fields, Fields)
        self.mainf = fields
        _lib.load()
        self._sfx = "" if fields.precision_float == "single" else "_f64"
        fields.append_instance(self)

    def update_e(self):
        f = self.mainf
        if f.ce_on and f.ce_mat is not None:
            _lib.call("mifdtd_update_e_lut", f.nx, f.ny, f.nz_pitch, f.pad, *f.ptrs(f.ehs),
                      *[m.data_ptr() for m in f.ce_mat], f.ce_lut.data_ptr(), _lib.stream_handle(f.stream))
            return
        ces = f.ces if f.ce_on else (None, None, None)
        _lib.call("mifdtd_update_e" + self._sfx, f.nx, f.ny, f.nz_pitch, f.pad, *f.ptrs(f.ehs), *f.ptrs(ces),
                  _lib.stream_handle(f.stream))

    def update_h(self):
        f = self.mainf
        chs = f.chs if f.ch_on else (None, None, None)
        _lib.call("mifdtd_update_h" + self._sfx, f.nx, f.ny, f.nz_pitch, f.pad, *f.ptrs(f.ehs), *f.ptrs(chs),
                  _lib.stream_handle(f.stream))


class CoreSplit(Core):
    """KEMP ``core_split.Core`` (``K7/cpu/core_split.py:8-123``): ``update_e(part)`` /
    ``update_h(part)`` update the flat element range of ``part`` -- '' the whole grid,
    'pre' / 'post' the ranges that file sets (E: planes 1-2 / plane 0; H: planes 0-1 /
    plane 2, in units of ny * nz_pitch elements) -- through ``mifdtd_update_*_range``."""

    def __init__(self, fields):
        super().__init__(fields)
        f = fields
        nyz = f.ny * f.nz_pitch
        self.e_ranges = {"": (0, f.nx * nyz), "pre": (nyz, 3 * nyz), "post": (0, nyz)}
        self.h_ranges = {"": (0, f.nx * nyz), "pre": (0, 2 * nyz), "post": (2 * nyz, 3 * nyz)}

    def update_e(self, part=""):
        f = self.mainf
        if f.ce_on and f.ce_mat is not None:
            raise ValueError("CoreSplit: the material-index CE form has no range entry point; "
                             "set CE arrays with set_ces")
        idx0, nmax = self.e_ranges[part]
        ces = f.ces if f.ce_on else (None, None, None)
        _lib.call("mifdtd_update_e" + self._sfx + "_range", f.nx, f.ny, f.nz_pitch, f.pad, idx0, nmax,
                  *f.ptrs(f.ehs), *f.ptrs(ces), _lib.stream_handle(f.stream))

    def update_h(self, part=""):
        f = self.mainf
        idx0, nmax = self.h_ranges[part]
        chs = f.chs if f.ch_on else (None, None, None)
        _lib.call("mifdtd_update_h" + self._sfx + "_range", f.nx, f.ny, f.nz_pitch, f.pad, idx0, nmax,
                  *f.ptrs(f.ehs), *f.ptrs(chs), _lib.stream_handle(f.stream))
