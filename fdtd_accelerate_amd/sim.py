"""Step-loop driver for one Yee + CPML simulation on the GPU.

``YeeSim`` assembles the KEMP-style objects (``Fields`` + ``Core`` + ``Cpml`` +
``IncidentDirect``) in the reference's order (``K7/main.py:296-326``,
``KU/cpu/test/020-dielectric-cpml.py:58-73``) and records probe time series on the
device (one gather per step into a device buffer, copied to the host once at the
end).  ``engine='kemp'`` runs the in-place per-half-step kernels (the drop-in
surface); ``engine='fused'`` runs the single-pass fused step and ``engine='slab'``
the in-place slab-ordered whole step (see ``fused.py``).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .core import Core
from .cpml import Cpml
from .fields import Fields
from .incident import IncidentDirect

_NAMES = ("ex", "ey", "ez", "hx", "hy", "hz")


class YeeSim:
    def __init__(self, nx, ny, nz, npml=0, coeff_use="e", cpml_faces=("fb", "fb", "fb"),
                 cpml_R=1e-6, cpml_m=4, segment_nbytes=64, engine="kemp", fused_params=None):
        if engine not in ("kemp", "fused", "slab"):
            raise ValueError(f"engine must be 'kemp', 'fused' or 'slab' : {engine!r}")
        self.fields = Fields(nx, ny, nz, coeff_use=coeff_use, segment_nbytes=segment_nbytes)
        self.core = Core(self.fields)
        self.cpml = Cpml(self.fields, npml, cpml_faces, m=cpml_m, R=cpml_R) if npml > 0 else None
        self.sources = []
        self.probe_points = []
        self._probe_off = None
        self._series = []
        self.engine = engine
        self.tstep = 0
        self._fused = None
        self._slab = None
        self._fused_params = fused_params

    # -------------------------------------------------------------- set-up
    def set_ce(self, cex, cey, cez):
        self.fields.set_ces(cex, cey, cez)

    def add_source(self, str_f, pt0, pt1, tfunc, spatial_value=1.0, is_overwrite=False):
        src = IncidentDirect(self.fields, str_f, pt0, pt1, tfunc, spatial_value, is_overwrite)
        self.sources.append(src)
        return src

    def add_probe(self, str_f, pt):
        f = self.fields
        i, j, k = pt
        if not (0 <= i < f.nx and 0 <= j < f.ny and 0 <= k < f.nz):
            raise ValueError(f"probe point {pt} outside the grid {f.ns}")
        self.probe_points.append((_NAMES.index(str_f), (int(i), int(j), int(k))))
        self._probe_off = None

    def _probe_offsets(self):
        if self._probe_off is None and self.probe_points:
            f = self.fields
            offs = [(i * f.ny + j) * f.nz_pitch + k for _, (i, j, k) in self.probe_points]
            self._probe_off = torch.tensor(offs, dtype=torch.int64, device=f.device)
        return self._probe_off

    # -------------------------------------------------------------- stepping
    def _record(self, buf, t):
        """Gather every probe (insertion order) into row t of the device buffer."""
        offs = self._probe_offsets()
        f = self.fields
        for n, (c, _) in enumerate(self.probe_points):
            _lib.call("mifdtd_gather", f.ehs[c].data_ptr(), offs[n:n + 1].data_ptr(), 1,
                      buf[t, n:n + 1].data_ptr(), _lib.stream_handle(f.stream))

    def run(self, nsteps):
        """Advance ``nsteps`` steps; returns the probe series of these steps (host array
        (nsteps, nprobes), columns in the order the probes were added) or None."""
        f = self.fields
        nprobe = len(self.probe_points)
        buf = torch.zeros((nsteps, nprobe), dtype=torch.float32, device=f.device) if nprobe else None
        if self.engine == "fused":
            from .fused import FusedStepper
            if self._fused is None:
                self._fused = FusedStepper(self, **(self._fused_params or {}))
            self._fused.run(nsteps, buf)
        elif self.engine == "slab":
            from .fused import SlabStepper
            if self._slab is None:
                self._slab = SlabStepper(self, **(self._fused_params or {}))
            self._slab.run(nsteps, buf)
        else:
            for t in range(nsteps):
                f.update_e()
                f.update_h()
                if buf is not None:
                    self._record(buf, t)
        self.tstep += nsteps
        if buf is None:
            return None
        s = buf.cpu().numpy()
        self._series.append(s)
        return s

    def series(self):
        return np.concatenate(self._series) if self._series else None

    def synchronize(self):
        self.fields.enqueue_barrier()

    def field(self, name):
        return self.fields.get(name)
