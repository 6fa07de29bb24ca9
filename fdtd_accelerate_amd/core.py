"""KEMP ``Core`` (``K7/cpu/core.py:8-123``, ``K7/gpu/core.py``): the Yee curl updates of a
``Fields`` through the HIP kernels of ``libmifdtd.so`` (``mifdtd_update_e/h``)."""
from __future__ import annotations

from . import _lib, util
from .fields import Fields


class Core:
    priority_type = "core"

    def __init__(self, fields):
        util.check_type("fields", fields, Fields)
        self.mainf = fields
        _lib.load()
        fields.append_instance(self)

    def update_e(self):
        f = self.mainf
        if f.ce_on and f.ce_mat is not None:
            _lib.call("mifdtd_update_e_lut", f.nx, f.ny, f.nz_pitch, f.pad, *f.ptrs(f.ehs),
                      *[m.data_ptr() for m in f.ce_mat], f.ce_lut.data_ptr(), _lib.stream_handle(f.stream))
            return
        ces = f.ces if f.ce_on else (None, None, None)
        _lib.call("mifdtd_update_e", f.nx, f.ny, f.nz_pitch, f.pad, *f.ptrs(f.ehs), *f.ptrs(ces),
                  _lib.stream_handle(f.stream))

    def update_h(self):
        f = self.mainf
        chs = f.chs if f.ch_on else (None, None, None)
        _lib.call("mifdtd_update_h", f.nx, f.ny, f.nz_pitch, f.pad, *f.ptrs(f.ehs), *f.ptrs(chs),
                  _lib.stream_handle(f.stream))
