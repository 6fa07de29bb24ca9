"""Total-field / scattered-field plane-wave source, normal incidence (kuFDTD ``Tfsf``,
``kuFDTD/v0.2.0/dim3/tfsf.py:90-242``).

A float32 1-D line (``Fdtd1d``, ``tfsf.py:14-40``) carries the incident wave along the
propagation axis: a hard source E[3] = sin(w dt tstep), two-step-delay ends, and the
same half-step updates as the 3-D grid in vacuum (CE = CH = 0.5).  Line point 5 is the
box's first plane.  After each 3-D half step (priority 'incident') the E components on
the planes just outside the box faces, and the H components on the faces, take the
incident correction 0.5 * c * (line value), c = +-sin / cos of the polarization angle,
per the face and coefficient tables of ``calc_face_list`` / ``calc_inc_list``
(``tfsf.py:138-226``).  The box [pt0, pt1] is inclusive; None takes 1 / n-2 as in
``tfsf.py:104-108``.  Units are the grid's: dx = 1, light velocity 1, dt = fields.dt,
wavelength in cells.  The reference's oblique branch is empty; only 'normal' exists.
The in-place engine runs it from the instance list; the fused engine (one plane wave)
applies the patches in its sweep and shell passes and advances the line around each step.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib, util
from .fields import Fields


def _box(sl):
    """Inclusive (i0, i1, j0, j1, k0, k1) of a 3-tuple of ints / slices."""
    out = []
    for v in sl:
        if isinstance(v, slice):
            out += [v.start, v.stop - 1]
        else:
            out += [v, v]
    return out


class Tfsf:
    priority_type = "incident"

    def __init__(self, fields, pt0, pt1, apply_direction, wavelength, propagation_direction=("normal", "x"),
                 polarization_angle=0.0):
        util.check_type("fields", fields, Fields)
        fields.require_single("Tfsf")
        util.check_type("apply_direction", apply_direction, (list, tuple), str)
        if len(apply_direction) != 3 or any(set(d) - set("fb") for d in apply_direction):
            raise ValueError(f"argument 'apply_direction' must be three strings of 'f'/'b' : {apply_direction!r}")
        if propagation_direction[0] != "normal":
            raise NotImplementedError("only normal incidence exists (the reference's oblique branch is empty)")
        axis = propagation_direction[1]
        util.check_value("propagation axis", axis, ("x", "y", "z"))
        if not wavelength > 0:
            raise ValueError(f"wavelength must be positive : {wavelength}")
        ns = fields.ns
        p0 = [1 if v is None else int(v) for v in pt0]
        p1 = [ns[a] - 2 if v is None else int(v) for a, v in enumerate(pt1)]
        for a in range(3):
            if not 1 <= p0[a] <= p1[a] <= ns[a] - 2:
                raise ValueError(f"TFSF box {p0}..{p1} needs one cell of margin inside {tuple(ns)}")
        self.mainf = fields
        self.axis = "xyz".index(axis)
        self.pt0, self.pt1 = p0, [v + 1 for v in p1]  # tfsf.py:111 (pt2 += 1)
        self.wdt = 2.0 * math.pi * fields.dt / wavelength
        self.tstep = 1
        n = self.pt1[self.axis] - self.pt0[self.axis] + 1
        self.n_line = n + 10
        dev = fields.device
        self.line_e = torch.zeros(self.n_line, dtype=torch.float32, device=dev)
        self.line_h = torch.zeros(self.n_line, dtype=torch.float32, device=dev)
        self.abc = torch.zeros(4, dtype=torch.float32, device=dev)
        self.patches_e, self.patches_h = self._patches(apply_direction, float(polarization_angle))
        fields.append_instance(self)

    def _patches(self, apply_direction, p_angle):
        a = self.axis
        i1d, i2d = 5, self.pt1[a] - self.pt0[a] + 5
        si, co = float(np.float32(np.sin(p_angle))), float(np.float32(np.cos(p_angle)))
        # line start index of the slices sl1..sl5 of calc_inc_list
        sl1, sl2, sl3, sl4, sl5 = i1d - 1, i1d, i1d, i1d, i2d
        # Inc1..Inc6: (ce1, ce2, ch1, ch2, start of sliceE1, sliceE2, sliceH1, sliceH2)
        inc = [(-si, co, co, si, sl1, sl1, sl2, sl2), (si, -co, -co, -si, sl5, sl5, sl5, sl5),
               (0.0, si, 0.0, -co, sl4, sl3, sl3, sl4), (0.0, -si, 0.0, co, sl4, sl3, sl3, sl4),
               (-co, 0.0, -si, 0.0, sl3, sl4, sl4, sl3), (co, 0.0, si, 0.0, sl3, sl4, sl4, sl3)]
        order = {0: [0, 1, 2, 3, 4, 5], 1: [4, 5, 0, 1, 2, 3], 2: [2, 3, 4, 5, 0, 1]}[a]
        (i1, j1, k1), (i2, j2, k2) = self.pt0, self.pt1
        slx, slxp, sly, slyp, slz, slzp = (slice(i1, i2), slice(i1, i2 + 1), slice(j1, j2), slice(j1, j2 + 1),
                                           slice(k1, k2), slice(k1, k2 + 1))
        # per face: (e1, e2, h1, h2 components, slice_e1, slice_e2, slice_h1, slice_h2)
        faces = [((1, 2, 4, 5), (i1 - 1, slyp, slz), (i1 - 1, sly, slzp), (i1, sly, slzp), (i1, slyp, slz)),
                 ((1, 2, 4, 5), (i2, slyp, slz), (i2, sly, slzp), (i2, sly, slzp), (i2, slyp, slz)),
                 ((2, 0, 5, 3), (slx, j1 - 1, slzp), (slxp, j1 - 1, slz), (slxp, j1, slz), (slx, j1, slzp)),
                 ((2, 0, 5, 3), (slx, j2, slzp), (slxp, j2, slz), (slxp, j2, slz), (slx, j2, slzp)),
                 ((0, 1, 3, 4), (slxp, sly, k1 - 1), (slx, slyp, k1 - 1), (slx, slyp, k1), (slxp, sly, k1)),
                 ((0, 1, 3, 4), (slxp, sly, k2), (slx, slyp, k2), (slx, slyp, k2), (slxp, sly, k2))]
        pe, ph = _lib.TfsfPatches(), _lib.TfsfPatches()
        pe.axis = ph.axis = a
        for f in range(6):
            if ("f" if f % 2 == 0 else "b") not in apply_direction[f // 2]:
                continue
            comps, se1, se2, sh1, sh2 = faces[f]
            ce1, ce2, ch1, ch2, bE1, bE2, bH1, bH2 = inc[order[f]]
            for P, comp, sl, coef, base in ((pe, comps[0], se1, ch1, bH1), (pe, comps[1], se2, ch2, bH2),
                                            (ph, comps[2], sh1, ce1, bE1), (ph, comps[3], sh2, ce2, bE2)):
                q = P.n
                P.comp[q] = comp
                for m, v in enumerate(_box(sl)):
                    P.box[q][m] = v
                P.base[q] = base
                P.coef[q] = coef
                P.n = q + 1
        return pe, ph

    def _faces(self, P, line):
        f = self.mainf
        ptrs = (ctypes.c_void_p * 6)(*[t.data_ptr() for t in f.ehs])
        _lib.call("mifdtd_tfsf_faces", ctypes.byref(P), ptrs, line.data_ptr(), self.n_line, f.nx, f.ny, f.nz_pitch,
                  _lib.stream_handle(f.stream))

    def update_e(self):
        f = self.mainf
        src = float(np.float32(np.sin(self.wdt * self.tstep)))  # tfsf.py:38
        self.tstep += 1
        _lib.call("mifdtd_tfsf_line", self.line_e.data_ptr(), self.line_h.data_ptr(), self.abc.data_ptr(),
                  self.n_line, 1, src, _lib.stream_handle(f.stream))
        self._faces(self.patches_e, self.line_h)

    def update_h(self):
        f = self.mainf
        _lib.call("mifdtd_tfsf_line", self.line_e.data_ptr(), self.line_h.data_ptr(), self.abc.data_ptr(),
                  self.n_line, 0, 0.0, _lib.stream_handle(f.stream))
        self._faces(self.patches_h, self.line_e)
