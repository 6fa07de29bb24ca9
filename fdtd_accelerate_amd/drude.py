"""Drude dispersion by the auxiliary-current method (kuFDTD ``Drude``,
``kuFDTD/v0.2.0/dim3/cpu/matter.py:104-221`` with ``cpu/core/drude_ade.c``).

On the box [pt0, pt1] (inclusive), right after the core E update (priority 'current',
before CPML-E and the sources), each E component c takes
    e' = e - CE_c J_c ;  J_c' = CJA_c J_c + CJB_c e'
with CJA = (2 - gamma dt) / (2 + gamma dt) and CJB = 4 (dt wp)^2 / (2 + gamma dt)
computed in float32 as ``matter.py:192-202`` does.  The plasma frequency wp and the
damping gamma (in the grid's units: dx = 1, dt = fields.dt) are scalars, region-sized
arrays, or a triple of them (one per component).  CE is the grid's E coefficient (0.5
without CE arrays).  Supported by the in-place engine and (one box) by the fused engine,
whose sweep and shell E pass apply it in the same order; the slab and batched engines
reject it.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib, util
from .fields import Fields


def _per_component(v, shape):
    if isinstance(v, (list, tuple)) and len(v) == 3:
        return [np.broadcast_to(np.asarray(x, np.float32), shape) for x in v]
    a = np.broadcast_to(np.asarray(v, np.float32), shape)
    return [a, a, a]


class Drude:
    priority_type = "current"

    def __init__(self, fields, pt0, pt1, plasma_freq, gamma=0.0):
        util.check_type("fields", fields, Fields)
        fields.require_single("Drude")
        p0, p1 = [int(v) for v in pt0], [int(v) for v in pt1]
        for a, n in enumerate(fields.ns):
            if not 0 <= p0[a] <= p1[a] < n:
                raise ValueError(f"Drude region {p0}..{p1} outside the grid {tuple(fields.ns)}")
        self.mainf = fields
        self.pt0, self.pt1 = p0, p1
        shape = tuple(b - a + 1 for a, b in zip(p0, p1))
        self.shape = shape
        dt = fields.dt
        cja, cjb = [], []
        for w, g in zip(_per_component(plasma_freq, shape), _per_component(gamma, shape)):
            tmp = 2 + g * dt                      # matter.py:192-194
            cja.append((2 - g * dt) / tmp)
            cjb.append(4 * (dt * w) ** 2 / tmp)
        dev = fields.device
        self.cja = [torch.from_numpy(np.ascontiguousarray(c, np.float32)).to(dev) for c in cja]
        self.cjb = [torch.from_numpy(np.ascontiguousarray(c, np.float32)).to(dev) for c in cjb]
        self.J = [torch.zeros(shape, dtype=torch.float32, device=dev) for _ in range(3)]
        fields.append_instance(self)

    def update_e(self):
        f = self.mainf
        P3 = ctypes.c_void_p * 3
        ce = P3(*[c.data_ptr() for c in f.ces]) if f.ce_on else None
        li, lj, lk = self.shape
        _lib.call("mifdtd_drude_e", P3(*[t.data_ptr() for t in f.ehs[:3]]), ce,
                  P3(*[t.data_ptr() for t in self.J]), P3(*[t.data_ptr() for t in self.cja]),
                  P3(*[t.data_ptr() for t in self.cjb]), f.nx, f.ny, f.nz_pitch, *self.pt0, li, lj, lk,
                  _lib.stream_handle(f.stream))

    def update_h(self):
        pass
