"""Batched engine: many independent ``YeeSim`` instances advanced together.

``BatchedYee(sims)`` uploads one descriptor per instance (``mifdtd_batch_create``) and
runs K steps as 2K launches for the whole batch (``mifdtd_batch_run``): the E pass of
every instance, then the H pass, with each instance's CPML, sources and probes fused
in (per-cell arithmetic of the in-place engine).  Source values are computed on the
host for all K steps and copied to the device once; probe series come back per instance.

``make_uneven_batch`` builds the uneven batch of the batched benchmark (64 cubes with edges
drawn from {64, 96, 128, 160, 192}, own sphere permittivity maps, 8-cell CPML, seeded
source and probe), and ``tune`` picks the launch options (graph capture, XCD block
runs) by timing them on the batch itself.
"""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np
import torch

from . import _lib, geometry
from .fused import in_place_only
from .sim import YeeSim
from .sources import gaussian_pulse

_COMP = {"ex": 0, "ey": 1, "ez": 2, "hx": 3, "hy": 4, "hz": 5}
_TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_batch.json")


def instance_descs(sims):
    """``mifdtd_instance`` descriptors of the simulations (a ctypes array) and the shared
    CPML tables (bE, aE, bH, aH device pointers, or Nones without CPML)."""
    ref = next((s.cpml for s in sims if s.cpml is not None), None)
    # CE as material indices when every instance has them (compress_ce), else the arrays
    tables = all(s.fields.ce_on and s.fields.ce_mat is not None for s in sims)
    descs = (_lib.InstanceDesc * len(sims))()
    for n, s in enumerate(sims):
        f = s.fields
        d = descs[n]
        d.nx, d.ny, d.nz_pitch, d.pad = f.nx, f.ny, f.nz_pitch, f.pad
        for c in range(6):
            d.e[c] = f.ehs[c].data_ptr()
        if f.ce_on:
            for c in range(3):
                d.ce[c] = f.ces[c].data_ptr()
        if tables:
            for c in range(3):
                d.ce_mat[c] = f.ce_mat[c].data_ptr()
            d.ce_lut = f.ce_lut.data_ptr()
        if s.cpml is not None:
            d.npml = s.cpml.npml
            d.psiE = s.cpml.psiE_buf.data_ptr()
            d.psiH = s.cpml.psiH_buf.data_ptr()
        d.nsrc = len(s.sources)
        for k, src in enumerate(s.sources):
            d.src_comp[k] = _COMP[src.str_f]
            box = (src.pt0[0], src.pt1[0], src.pt0[1], src.pt1[1], src.pt0[2], src.pt1[2])
            for m in range(6):
                d.src_box[k][m] = box[m]
            d.src_sv[k] = None if src.sv is None else src.sv.data_ptr()
            d.src_svs[k] = src.svs
            d.src_overwrite[k] = int(src.is_overwrite)
        d.nprobe = len(s.probe_points)
        for k, (c, (i, j, kk)) in enumerate(s.probe_points):
            d.probe_comp[k] = c
            d.probe_ijk[k][0], d.probe_ijk[k][1], d.probe_ijk[k][2] = i, j, kk
    tabs = (None,) * 4 if ref is None else (ref.bE.data_ptr(), ref.aE.data_ptr(), ref.bH.data_ptr(), ref.aH.data_ptr())
    return descs, tabs


def source_values(sims, nsteps, device):
    """Device (nsteps, sum of sources) float32 values: column n is the n-th source of the
    batch (instances in order), row k its tfunc(tstep + k) cast to float32."""
    total = sum(len(s.sources) for s in sims)
    tv = np.zeros((max(nsteps, 1), max(total, 1)), np.float32)
    col = 0
    for s in sims:
        for src in s.sources:
            for k in range(nsteps):
                tv[k, col] = np.float32(src.tfunc(src.tstep + k))
            col += 1
    return torch.from_numpy(tv).to(device)


def advance(sims, nsteps):
    """Move the sources' and simulations' step counters on after nsteps batched steps."""
    for s in sims:
        for src in s.sources:
            src.tstep += nsteps
        s.tstep += nsteps


class BatchedYee:
    kernel_name = "k_batch (E pass + H pass of every instance)"

    def __init__(self, sims, use_graph=None, remap_rows=None, stream=None):
        if not sims:
            raise ValueError("need at least one simulation")
        self.sims = list(sims)
        self.device = sims[0].fields.device
        self.stream = stream if stream is not None else torch.cuda.Stream(self.device)
        self._h = None
        self._create()
        key = self.key()
        tuned = _load_table().get(key, {})
        self.use_graph = bool(tuned.get("use_graph", False)) if use_graph is None else bool(use_graph)
        self.remap_rows = int(tuned.get("remap_rows", 0)) if remap_rows is None else int(remap_rows)

    def _state(self):
        """What the device descriptors hold: every buffer address, the source and probe
        counts and the update instances of each simulation."""
        out = []
        for s in self.sims:
            f = s.fields
            out += [t.data_ptr() for t in f.ehs]
            out += [t.data_ptr() for t in f.ces] if f.ce_on else []
            out += [t.data_ptr() for t in f.ce_mat] + [f.ce_lut.data_ptr()] if f.ce_mat is not None else [0]
            out += [s.cpml.psiE_buf.data_ptr(), s.cpml.psiH_buf.data_ptr()] if s.cpml is not None else [0]
            out += [0 if src.sv is None else src.sv.data_ptr() for src in s.sources]
            out += [len(s.sources), len(s.probe_points), tuple(id(i) for i in f.instance_list)]
        return tuple(out)

    def _refresh(self):
        """Build the device descriptors again when a simulation's buffers, sources,
        probes or update instances changed since they were built (another engine may have
        moved the field arrays or swapped the psi buffers)."""
        if self._state() != self._built:
            self.close()
            self._create()

    def _create(self):
        sims = self.sims
        npmls = {s.cpml.npml for s in sims if s.cpml is not None}
        if len(npmls) > 1:
            raise ValueError(f"all instances must share one npml : {sorted(npmls)}")
        # one set of b/a tables serves the whole batch: every instance's CPML (grading m,
        # R, kappa, alpha, dt) must give the same float32 tables
        tabs = [s.cpml for s in sims if s.cpml is not None]
        for cp in tabs[1:]:
            if any(not torch.equal(x, y) for x, y in zip((cp.bE, cp.aE, cp.bH, cp.aH),
                                                          (tabs[0].bE, tabs[0].aE, tabs[0].bH, tabs[0].aH))):
                raise ValueError("all instances must share the same CPML tables (grading, R, kappa, alpha)")
        ce = {s.fields.ce_on for s in sims}
        if len(ce) > 1:
            raise ValueError("CE arrays must be used by every instance or by none")
        for s in sims:
            if s.fields.ch_on:
                raise ValueError("the batched engine uses CH = 0.5")
            if s.cpml is not None and any(d != "fb" and set(d) != {"f", "b"} for d in s.cpml.apply_direction):
                raise ValueError("the batched engine needs CPML on all six faces (or none)")
            if len(s.sources) > _lib.MAX_SRC or len(s.probe_points) > _lib.MAX_PROBE:
                raise ValueError("too many sources or probes on an instance")
            for cls, what in in_place_only():
                if any(isinstance(i, cls) for i in s.fields.instance_list):
                    raise ValueError(f"the batched engine has no {what}")
        descs, tabs = instance_descs(sims)
        h = ctypes.c_void_p()
        _lib.call("mifdtd_batch_create", descs, len(sims), *tabs, ctypes.byref(h))
        self._h = h
        ns, npb, nb = ctypes.c_int(), ctypes.c_int(), ctypes.c_uint()
        _lib.call("mifdtd_batch_sizes", h, ctypes.byref(ns), ctypes.byref(npb), ctypes.byref(nb))
        self.total_src, self.total_probe, self.total_blocks = ns.value, npb.value, nb.value
        self.cells = sum(s.fields.nx * s.fields.ny * s.fields.nz for s in sims)
        self._built = self._state()

    def key(self):
        return json.dumps(sorted((s.fields.nx, s.fields.ny, s.fields.nz) for s in self.sims) +
                          [self.sims[0].cpml.npml if self.sims[0].cpml else 0])

    def close(self):
        if getattr(self, "_h", None):
            _lib.call("mifdtd_batch_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _tvals(self, nsteps):
        return source_values(self.sims, nsteps, self.device)

    def run(self, nsteps, collect=True):
        """Advance every instance nsteps steps; returns the per-instance probe series
        (list of host arrays) when ``collect``."""
        self._refresh()
        tv = self._tvals(nsteps)
        buf = torch.zeros((max(nsteps, 1), max(self.total_probe, 1)), dtype=torch.float32, device=self.device)
        # the batch stream runs after the work already queued on the instances' stream
        self.stream.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(self.stream):
            _lib.call("mifdtd_batch_run", self._h, nsteps, tv.data_ptr(), buf.data_ptr(), int(self.use_graph),
                      self.remap_rows, self.stream.cuda_stream)
        torch.cuda.current_stream(self.device).wait_stream(self.stream)
        tv.record_stream(self.stream)
        buf.record_stream(self.stream)
        advance(self.sims, nsteps)
        if not collect:
            return None
        host = buf.cpu().numpy()[:nsteps]
        out, col = [], 0
        for s in self.sims:
            k = len(s.probe_points)
            part = host[:, col:col + k].copy() if k else None
            if k:
                s._series.append(part)
            out.append(part)
            col += k
        return out

    def time_steps(self, nsteps):
        """Average ms per step of nsteps steps (events on the batch stream)."""
        self._refresh()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        tv = self._tvals(nsteps)
        buf = torch.zeros((max(nsteps, 1), max(self.total_probe, 1)), dtype=torch.float32, device=self.device)
        self.stream.wait_stream(torch.cuda.current_stream(self.device))
        torch.cuda.synchronize()
        a.record(self.stream)
        _lib.call("mifdtd_batch_run", self._h, nsteps, tv.data_ptr(), buf.data_ptr(), int(self.use_graph),
                  self.remap_rows, self.stream.cuda_stream)
        b.record(self.stream)
        torch.cuda.synchronize()
        advance(self.sims, nsteps)
        return a.elapsed_time(b) / max(nsteps, 1)


def _load_table():
    try:
        with open(_TABLE_PATH) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def tune(batch, candidates=((False, 0), (True, 0), (False, 4), (True, 4), (False, 16), (True, 16)), steps=10,
         save=True):
    """Time each (use_graph, remap_rows) on the batch (advancing it) and keep the best."""
    if not candidates:
        raise ValueError("no (use_graph, remap_rows) candidates to time")
    res = {}
    for g, r in candidates:
        batch.use_graph, batch.remap_rows = g, r
        batch.time_steps(2)
        res[(g, r)] = batch.time_steps(steps)
    best = min(res, key=res.get)
    batch.use_graph, batch.remap_rows = best
    if save:
        t = _load_table()
        t[batch.key()] = {"use_graph": bool(best[0]), "remap_rows": int(best[1])}
        with open(_TABLE_PATH, "w") as fh:
            json.dump(t, fh, indent=1)
    return best, res


def make_uneven_batch(seed=3, count=64, sizes=(64, 96, 128, 160, 192), npml=8, engine="kemp", ce_table=True):
    """The uneven batch: ``count`` cubes with edges drawn uniformly from ``sizes``, each
    with its own 10-40 non-overlapping spheres (eps_r U[1,12], radius U[3, N/8]), CE on,
    CPML ``npml`` (m = 4, R = 1e-6), an Ez source at a seeded interior point driven by the
    Gaussian-modulated pulse, and an Ez probe 8 cells away along x.  With ``ce_table``
    each instance also holds its CE as uint8 material indices + table (compress_ce; the
    same values, read as 3 B instead of 12 B per cell)."""
    rng = np.random.default_rng(seed)
    sims = []
    for n in range(count):
        N = int(rng.choice(sizes))
        sim = YeeSim(N, N, N, npml=npml, coeff_use="e", engine=engine)
        nsph = int(rng.integers(10, 41))
        lo, hi = (npml + 1,) * 3, (N - npml - 1,) * 3
        sph = geometry.random_spheres(int(rng.integers(1 << 30)), nsph, (3.0, max(4.0, N / 8)), (1.0, 12.0), lo, hi)
        geometry.paint_ce(sim.fields.ces, sph, (N, N, N))
        if ce_table and not sim.fields.compress_ce():
            raise RuntimeError(f"instance {n}: more than 256 CE values")
        lo_s = npml + 2
        src = tuple(int(v) for v in rng.integers(lo_s, max(lo_s + 1, N - npml - 10), 3))
        sim.add_source("ez", src, src, gaussian_pulse())
        sim.add_probe("ez", (src[0] + 8, src[1], src[2]))
        sims.append(sim)
    torch.cuda.synchronize()
    return sims
