"""Convolutional PML on the six faces of a ``Fields`` (kuFDTD ``CpmlNonKappa``,
``KU/cpu/cpml.py:16-203``), driven through ``mifdtd_cpml_e/h``.

Coefficients (``KU/cpu/cpml.py:107-126`` in normalised units, eps0 = 1, dx = 1):
  sigma_max = -(m+1) ln(R) / (2 eta npml dx)          (eta = 1; polynomial grading m)
  sigma_E   front[p+1] = smax((npml-0.5-p)/npml)^m,  back[npml+1+p] = smax((0.5+p)/npml)^m
  sigma_H   front[p+1] = smax((npml-p)/npml)^m,      back[npml+2+p] = smax((1+p)/npml)^m
  b = exp(-(sigma/kappa + alpha) dt),  a = sigma/(sigma kappa + alpha kappa^2) (b - 1)
with kappa = 1 and alpha = 0 by default (a = 0 where sigma = 0); computed in float64 on
the host and stored as float32 device tables of 2(npml+1) entries.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import _lib, util
from .fields import Fields


def sigma_max(npml, m=4, R=1e-6, eta=1.0, dx=1.0):
    return -(m + 1) * math.log(R) / (2.0 * eta * npml * dx)


def cpml_coefficients(npml, m=4, R=1e-6, dt=0.5, kappa=1.0, alpha=0.0):
    """float64 (bE, aE, bH, aH) tables in kuFDTD's 2(npml+1) layout."""
    smax = sigma_max(npml, m, R)
    sE = np.zeros(2 * (npml + 1))
    sH = np.zeros(2 * (npml + 1))
    p = np.arange(npml)
    sE[p + 1] = smax * ((npml - 0.5 - p) / npml) ** m
    sE[npml + 1 + p] = smax * ((0.5 + p) / npml) ** m
    sH[p + 1] = smax * ((npml - p) / npml) ** m
    sH[npml + 2 + p] = smax * ((1.0 + p) / npml) ** m

    def ba(s):
        b = np.exp(-(s / kappa + alpha) * dt)
        den = s * kappa + alpha * kappa ** 2
        a = np.where(den > 0, s / np.where(den > 0, den, 1.0) * (b - 1.0), 0.0)
        return b, a

    bE, aE = ba(sE)
    bH, aH = ba(sH)
    return bE, aE, bH, aH


def psi_shape(axis, nx, ny, nz_pitch, npml):
    """psi layout per face: x (npml, ny, nzp), y (nx, npml, nzp), z (nx, ny, npml)."""
    return [(npml, ny, nz_pitch), (nx, npml, nz_pitch), (nx, ny, npml)][axis]


class Cpml:
    """CPML of thickness ``npml`` on the faces named by ``apply_direction`` (one string
    of 'f'/'b' per axis, as kuFDTD's ``CpmlNonKappa(Npml, ('fb','fb','fb'))``)."""

    priority_type = "pml"

    def __init__(self, fields, npml, apply_direction=("fb", "fb", "fb"), m=4, R=1e-6,
                 kappa=1.0, alpha=0.0):
        util.check_type("fields", fields, Fields)
        fields.require_single("Cpml")
        util.check_type("npml", npml, int)
        util.check_type("apply_direction", apply_direction, (list, tuple), str)
        if len(apply_direction) != 3 or any(set(d) - set("fb") for d in apply_direction):
            raise ValueError(f"argument 'apply_direction' must be three strings of 'f'/'b' : {apply_direction!r}")
        if not 1 <= npml <= 64:
            raise ValueError(f"argument 'npml' must be in [1, 64] : {npml}")
        for n, ax in zip(fields.ns, "xyz"):
            if n < 2 * npml + 2:
                raise ValueError(f"grid extent {n} along {ax} is too small for npml {npml}")
        self.mainf = fields
        self.npml = npml
        self.apply_direction = tuple(apply_direction)
        dev = fields.device
        bE, aE, bH, aH = cpml_coefficients(npml, m, R, fields.dt, kappa, alpha)
        self.tables_f64 = (bE, aE, bH, aH)
        t = lambda x: torch.tensor(x, dtype=torch.float32, device=dev)
        self.bE, self.aE, self.bH, self.aH = t(bE), t(aE), t(bH), t(aH)
        nx, ny, nzp = fields.ns_pitch
        # psi of all six faces packed in one buffer per half (psi1 then psi2 of each face,
        # x faces, y faces, z faces); faces that are off keep their (zero) space
        self._sizes = [npml * ny * nzp, nx * npml * nzp, nx * ny * npml]
        total = 4 * sum(self._sizes)
        self.psiE_buf = torch.zeros(total, dtype=torch.float32, device=dev)
        self.psiH_buf = torch.zeros(total, dtype=torch.float32, device=dev)
        self.psiE_alt_buf = None  # second psi-E buffer of the fused engine (allocated on demand)
        self._build_views()
        _lib.load()
        fields.append_instance(self)

    def face_offset(self, face, comp):
        """Float offset of psi<comp+1> of face (2*axis+side) in a packed buffer."""
        ax, sd = divmod(face, 2)
        sX, sY, sZ = self._sizes
        base = [0, 4 * sX, 4 * sX + 4 * sY][ax]
        return base + (2 * sd + comp) * self._sizes[ax]

    def _views(self, buf):
        nx, ny, nzp = self.mainf.ns_pitch
        out = {}
        for ax in range(3):
            for sd, c in enumerate("fb"):
                if c in self.apply_direction[ax]:
                    f = 2 * ax + sd
                    n = self._sizes[ax]
                    shp = psi_shape(ax, nx, ny, nzp, self.npml)
                    o1, o2 = self.face_offset(f, 0), self.face_offset(f, 1)
                    out[ax, sd] = (buf[o1:o1 + n].view(shp), buf[o2:o2 + n].view(shp))
        return out

    def _build_views(self):
        self.psiE = self._views(self.psiE_buf)
        self.psiH = self._views(self.psiH_buf)
        self._descE = self._desc(self.psiE, self.bE, self.aE)
        self._descH = self._desc(self.psiH, self.bH, self.aH)

    def alt_psiE(self):
        if self.psiE_alt_buf is None:
            self.psiE_alt_buf = torch.zeros_like(self.psiE_buf)
        return self.psiE_alt_buf

    def swap_psiE(self):
        """Make the alternate psi-E buffer current (after an odd number of fused steps)."""
        self.psiE_buf, self.psiE_alt_buf = self.psiE_alt_buf, self.psiE_buf
        self._build_views()

    def _desc(self, psi, b, a):
        d = _lib.CpmlHalf()
        d.npml = self.npml
        d.b = b.data_ptr()
        d.a = a.data_ptr()
        for (ax, sd), (p1, p2) in psi.items():
            d.psi1[2 * ax + sd] = p1.data_ptr()
            d.psi2[2 * ax + sd] = p2.data_ptr()
        return d

    def reset(self):
        self.psiE_buf.zero_()
        self.psiH_buf.zero_()
        if self.psiE_alt_buf is not None:
            self.psiE_alt_buf.zero_()

    def update_e(self):
        f = self.mainf
        ces = f.ces if f.ce_on else (None, None, None)
        _lib.call("mifdtd_cpml_e", f.nx, f.ny, f.nz_pitch, f.pad, self._descE, *f.ptrs(f.ehs),
                  *f.ptrs(ces), _lib.stream_handle(f.stream))

    def update_h(self):
        f = self.mainf
        chs = f.chs if f.ch_on else (None, None, None)
        _lib.call("mifdtd_cpml_h", f.nx, f.ny, f.nz_pitch, f.pad, self._descH, *f.ptrs(f.ehs),
                  *f.ptrs(chs), _lib.stream_handle(f.stream))


_FACE_E = {0: ("ey", "ez", "hy", "hz", 1, 2), 1: ("ex", "ez", "hx", "hz", 0, 2), 2: ("ex", "ey", "hx", "hy", 0, 1)}


def check_face_psi(fields, axis, npml, psi1, psi2):
    """psi of one face must have this engine's shape (``psi_shape``), not kuFDTD's
    (Npml, Ny-1, Nz), (Nx-1, Npml, Nz), (Nx-1, Ny-1, Npml+1) (``KU/cpu/cpml.py:35-85``),
    which the face kernels would index past its end."""
    nx, ny, nzp = fields.ns_pitch
    want = psi_shape(axis, nx, ny, nzp, npml)
    for name, p in (("psi1", psi1), ("psi2", psi2)):
        if not isinstance(p, torch.Tensor):
            raise TypeError(f"argument '{name}' type must be a torch.Tensor : {type(p)} is given")
        if tuple(p.shape) != want or p.dtype != torch.float32 or not p.is_contiguous():
            raise ValueError(f"argument '{name}' must be a contiguous float32 tensor of shape {want} "
                             f"(x: (npml, ny, nz_pitch), y: (nx, npml, nz_pitch), z: (nx, ny, npml)) : "
                             f"{tuple(p.shape)} {p.dtype} is given")


def face_update(fields, half, axis, side, npml, psi1, psi2, b, a):
    """One CPML face of one half step through ``mifdtd_cpml_face_<x|y|z>_<e|h>``
    (kuFDTD's per-face call, ``KU/cpu/cpml.py:131-203``, on this engine's layout)."""
    util.check_value("half", half, ("e", "h"))
    util.check_value("axis", axis, (0, 1, 2))
    util.check_value("side", side, (0, 1))
    check_face_psi(fields, axis, npml, psi1, psi2)
    f1, f2, g1, g2, c1, c2 = _FACE_E[axis]
    g = fields.get
    args = [fields.nx, fields.ny, fields.nz_pitch, fields.pad, npml, side]
    if half == "e":
        ces = (fields.ces[c1].data_ptr(), fields.ces[c2].data_ptr()) if fields.ce_on else (None, None)
        args += [g(f1).data_ptr(), g(f2).data_ptr(), g(g1).data_ptr(), g(g2).data_ptr(), *ces]
    else:
        args += [g(f1).data_ptr(), g(f2).data_ptr(), g(g1).data_ptr(), g(g2).data_ptr()]
    args += [psi1.data_ptr(), psi2.data_ptr(), b.data_ptr(), a.data_ptr(), _lib.stream_handle(fields.stream)]
    _lib.call(f"mifdtd_cpml_face_{'xyz'[axis]}_{half}", *args)
