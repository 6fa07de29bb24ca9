"""KEMP ``Pbc`` (``K7/cpu/pbc.py:5-39``): periodic boundaries along any of x, y, z.

After each E half step the E components tangential to a periodic axis take, on the
last plane of that axis, the values of the first plane; after each H half step the H
components take, on the first plane, the values of the last plane (the slice table of
``K7/util/common_exchange.py:1-29``).  The grid then has period n - 1 along the axis.
Priority 'pbc': after the core, CPML and source updates of the same half step.  Axes
are applied in the order given.  Supported by the in-place engine (``Fields`` instance
list) and by the fused engine (the copies after its E and H passes; its box leaves the
last plane of a periodic axis to the shell pass); the slab and batched engines reject it.
"""
from __future__ import annotations

from . import _lib, util
from .fields import Fields

# components tangential to each axis, per half step (indices into Fields.ehs)
_TANGENTIAL = {"x": {"e": (1, 2), "h": (4, 5)}, "y": {"e": (0, 2), "h": (3, 5)},
               "z": {"e": (0, 1), "h": (3, 4)}}


class Pbc:
    priority_type = "pbc"

    def __init__(self, fields, axes):
        util.check_type("fields", fields, Fields)
        fields.require_single("Pbc")
        util.check_type("axes", axes, str)
        if not set(axes) & set("xyz") or set(axes) - set("xyz"):
            raise AssertionError(f"axes option is wrong: {axes!r} is given")
        self.mainf = fields
        self.axes = axes
        fields.append_instance(self)

    def _update(self, e_or_h):
        f = self.mainf
        for axis in self.axes:
            c1, c2 = _TANGENTIAL[axis][e_or_h]
            _lib.call("mifdtd_pbc", f.ehs[c1].data_ptr(), f.ehs[c2].data_ptr(), f.nx, f.ny, f.nz_pitch, f.pad,
                      "xyz".index(axis), int(e_or_h == "e"), _lib.stream_handle(f.stream))

    def update_e(self):
        self._update("e")

    def update_h(self):
        self._update("h")
