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

from . import _lib, util
from .core import Core
from .cpml import Cpml
from .fields import Fields
from .incident import IncidentDirect

_NAMES = ("ex", "ey", "ez", "hx", "hy", "hz")


class YeeSim:
    def __init__(self, nx, ny, nz, npml=0, coeff_use="e", cpml_faces=("fb", "fb", "fb"),
                 cpml_R=1e-6, cpml_m=4, segment_nbytes=64, engine="kemp", fused_params=None, max_tstep=None):
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
        self.max_tstep = max_tstep
        self.savef_list = []

    @classmethod
    def from_geometry_h5(cls, geometry_h5_path, max_tstep=None, **kwargs):
        """A simulation sized by a geometry file: ``coeff_use``, ``nx``, ``ny``, ``nz``
        read from its root attributes as ``K7/main.py:70-75`` (``Fdtd3d``) reads them."""
        from .snapshot import read_geometry_h5
        coeff_use, nx, ny, nz = read_geometry_h5(geometry_h5_path)
        return cls(nx, ny, nz, coeff_use=coeff_use, max_tstep=max_tstep, **kwargs)

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

    def set_savefields(self, str_f, pt0, pt1, tstep_range, dir_path, tag=0, max_tstep=None, **options):
        """Snapshot files of components of the box [pt0, pt1] at the steps ``tstep`` with
        ``tstep % tgap == 0`` and ``t0 <= tstep <= t1`` (``K7/main.py:239-293`` and its
        loop ``:315-317``): ``tstep_range = (t0, t1, tgap)`` with t0 and t1 converted by
        ``convert_index`` over max_tstep and t1 incremented by one, as that file does;
        a tag already in use becomes one more than the largest; ``divide_info.pkl`` is
        written to ``dir_path``. Unlike the reference, where the last call's range holds
        for every SaveFields, each keeps its own range. ``options`` go to ``SaveFields``
        (``gzip_level``, ``nslots``, ``nthreads``). Returns the ``SaveFields``."""
        from .snapshot import SaveFields, write_divide_info
        util.check_type("tstep_range", tstep_range, (list, tuple), int)
        util.check_type("dir_path", dir_path, str)
        util.check_type("tag", tag, int)
        tmax = max_tstep if max_tstep is not None else self.max_tstep
        if tmax is None:
            raise ValueError("set_savefields needs max_tstep (YeeSim(max_tstep=...) or the argument)")
        if len(tstep_range) != 3 or tstep_range[2] < 1:
            raise ValueError(f"tstep_range must be (t0, t1, tgap) with tgap >= 1 : {tstep_range}")
        f = self.fields
        p0 = util.convert_indices(f.ns, pt0)
        p1 = util.convert_indices(f.ns, pt1)
        t0 = util.convert_index(tmax, tstep_range[0])
        t1 = util.convert_index(tmax, tstep_range[1]) + 1
        tgap = tstep_range[2]
        tags = [sf.tag for sf in self.savef_list]
        tag = tag if tag not in tags else max(tags) + 1
        savef = SaveFields(f, str_f, p0, p1, dir_path, tag, tmax, **options)
        savef.tag, savef.trange = tag, (t0, t1, tgap)
        self.savef_list.append(savef)
        write_divide_info(dir_path, util.shape_two_points(p0, p1), p0, p1, f.nx, savef.str_fs, tmax, tgap)
        return savef

    def _next_save(self, limit):
        """Steps (1..limit) until the next snapshot step, or limit when none comes sooner."""
        k = limit
        for sf in self.savef_list:
            t0, t1, tgap = sf.trange
            m = max(self.tstep // tgap + 1, -(-t0 // tgap)) * tgap
            if m <= t1 and m - self.tstep < k:
                k = m - self.tstep
        return k

    def wait_savefields(self):
        """Block until every snapshot file is written."""
        for sf in self.savef_list:
            sf.wait()

    def _advance(self, nsteps, buf):
        f = self.fields
        # the whole-step engines take the update instances (CPML faces, periodic axes, a
        # Drude box, a TFSF source) when they are built: a changed instance list builds
        # them again (the state carries over)
        instances = tuple(id(i) for i in f.instance_list)
        if instances != getattr(self, "_stepper_instances", instances):
            self._fused = self._slab = None
        self._stepper_instances = instances
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

    def run(self, nsteps):
        """Advance ``nsteps`` steps; returns the probe series of these steps (host array
        (nsteps, nprobes), columns in the order the probes were added) or None. Snapshot
        steps of ``set_savefields`` end a chunk of steps and enqueue their snapshot."""
        f = self.fields
        nprobe = len(self.probe_points)
        buf = torch.zeros((nsteps, nprobe), dtype=torch.float32, device=f.device) if nprobe else None
        done = 0
        while done < nsteps:
            k = self._next_save(nsteps - done) if self.savef_list else nsteps - done
            self._advance(k, None if buf is None else buf[done:done + k])
            done += k
            self.tstep += k
            for sf in self.savef_list:
                t0, t1, tgap = sf.trange
                if self.tstep % tgap == 0 and t0 <= self.tstep <= t1:
                    sf.save_fields(self.tstep)
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
