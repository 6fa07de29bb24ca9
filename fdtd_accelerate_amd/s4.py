"""4th-order symplectic FDTD on the GPU (the reference's ``4th_order-Symplectic`` path).

Layout and stage loop of ``S4/PyCUDA/100-base.py``: fields are (nx, ny, nz) Fortran
order (x contiguous; here torch tensors of shape (nz, ny, nx)), coefficient arrays
hold S/(24 eps_r), and one step is

    for s in 0..3:  H(c_s), E(d_s), E sources at time tstep + s/5
    H(c_4)

(``100-base.py:124-141``; the closing H uses c_4 = c1 as the PyCUDA version does).
The stage updates are ``mifdtd_s4_update_h/e`` (``S4/Python-C/dielectric.c``
arithmetic); the optional CPML (``mifdtd_s4_cpml_h/e``, new: the reference has none)
runs after each stage with the tables of that stage's time step w_s*S.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib
from .cpml import cpml_coefficients

S4_C = (0.17399689146541, -0.12038504121430, 0.89277629949778, -0.12038504121430, 0.17399689146541)
S4_D = (0.62337932451322, -0.12337932451322, -0.12337932451322, 0.62337932451322, 0.0)


def amplification(theta, c=S4_C, d=S4_D):
    """One step of the integrator on the oscillator e' = w h, h' = -w e with
    theta = w*dt: the 2x2 matrix mapping (e, h) (float64)."""
    M = np.eye(2)
    for s in range(5):
        M = np.array([[1.0, 0.0], [-c[s] * theta, 1.0]]) @ M  # h -= c_s*theta*e
        if s < 4:
            M = np.array([[1.0, d[s] * theta], [0.0, 1.0]]) @ M  # e += d_s*theta*h
    return M


def stability_theta(c=S4_C, d=S4_D):
    """Largest theta = w*dt with |trace(M)| <= 2 (bisection on the first crossing)."""
    lo, hi = 0.1, 0.1
    while abs(np.trace(amplification(hi, c, d))) <= 2.0:
        lo, hi = hi, hi + 0.05
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        if abs(np.trace(amplification(mid, c, d))) <= 2.0:
            lo = mid
        else:
            hi = mid
    return lo


def courant_limit():
    """S_max of the 3-D scheme: theta_max / (sqrt(3) * 7/3), 7/3 being the largest
    eigenvalue magnitude of the (27, -1)/24 difference (unit cell, c = 1)."""
    return stability_theta() / (math.sqrt(3.0) * 7.0 / 3.0)


_NAMES = ("ex", "ey", "ez", "hx", "hy", "hz")


class S4Sim:
    """One 4th-order simulation.  ``S`` is the Courant number (dt = S, dx = c = 1);
    default 0.9 * S_max.  ``eps`` (optional, (nx, ny, nz) array) sets CE = S/(24 eps)
    at every E component (the reference samples one array for all three, 100-base.py:84-86;
    ``set_ce`` takes three).  CH is the constant S/24 unless ``set_ch`` is called."""

    def __init__(self, nx, ny, nz, S=None, npml=0, cpml_R=1e-6, cpml_m=4, device="cuda"):
        if nx % 4:
            raise ValueError(f"nx must be a multiple of 4 : {nx}")
        if npml and min(nx, ny, nz) < 2 * npml + 4:
            raise ValueError(f"grid {(nx, ny, nz)} too small for npml {npml}")
        if not torch.cuda.is_available():
            raise RuntimeError("S4Sim needs a HIP device: the 4th-order kernels have no CPU path")
        _lib.load()
        self.nx, self.ny, self.nz = nx, ny, nz
        self.S = float(S) if S is not None else 0.9 * courant_limit()
        self.npml = npml
        self.device = torch.device(device)
        self.stream = torch.cuda.current_stream(self.device)
        shape = (nz, ny, nx)  # memory == Fortran (nx, ny, nz)
        self.ehs = [torch.zeros(shape, dtype=torch.float32, device=self.device) for _ in range(6)]
        self.c0 = float(np.float32(self.S / 24.0))
        self.ces = None
        self.ce_mat = self.ce_lut = None  # CE as uint8 material indices + table (set_ce)
        self.chs = None
        self.cl = [float(np.float32(c)) for c in S4_C]
        self.dl = [float(np.float32(d)) for d in S4_D]
        self.sources = []
        self.probe_points = []
        self._series = []
        self.tstep = 1
        if npml:
            self._init_cpml(cpml_R, cpml_m)

    # ---------------------------------------------------------------- set-up
    def _init_cpml(self, R, m):
        npml, nx, ny, nz = self.npml, self.nx, self.ny, self.nz
        dev = self.device

        def tab(w):
            bE, aE, bH, aH = cpml_coefficients(npml, m=m, R=R, dt=w * self.S)
            f = lambda t: torch.tensor(np.asarray(t, np.float32), device=dev)
            return f(bE), f(aE), f(bH), f(aH)

        self._tabE = [tab(S4_D[s])[:2] for s in range(4)]
        self._tabH = [tab(S4_C[s])[2:] for s in range(5)]
        sizes = [(nz, ny, npml), (nz, npml, nx), (npml, ny, nx)]  # torch shapes of F-order psi
        self.psiE = {}
        self.psiH = {}
        for ax in range(3):
            for sd in range(2):
                self.psiE[ax, sd] = tuple(torch.zeros(sizes[ax], device=dev) for _ in range(2))
                self.psiH[ax, sd] = tuple(torch.zeros(sizes[ax], device=dev) for _ in range(2))

    def _half(self, psi, b, a):
        d = _lib.CpmlHalf()
        d.npml = self.npml
        d.b, d.a = b.data_ptr(), a.data_ptr()
        for (ax, sd), (p1, p2) in psi.items():
            d.psi1[2 * ax + sd] = p1.data_ptr()
            d.psi2[2 * ax + sd] = p2.data_ptr()
        return d

    def _fortran(self, a):
        a = torch.as_tensor(np.asarray(a, np.float32) if not isinstance(a, torch.Tensor) else a)
        if tuple(a.shape) != (self.nx, self.ny, self.nz):
            raise ValueError(f"array shape {tuple(a.shape)} != {(self.nx, self.ny, self.nz)}")
        return a.to(self.device, torch.float32).permute(2, 1, 0).contiguous()

    def set_eps(self, eps):
        ce = torch.as_tensor(np.float32(self.S / 24.0) / np.asarray(eps, np.float32))
        self.set_ce(ce, ce, ce)

    def set_ce(self, cex, cey, cez):
        """CE arrays; when each component holds at most 256 distinct values the E stages
        read them as uint8 material indices into an exact table (39 instead of 48 bytes
        per cell, the same values; the CPML keeps the float arrays)."""
        self.ces = [self._fortran(c) for c in (cex, cey, cez)]
        self.ce_mat = self.ce_lut = None
        lut = torch.zeros((3, 256), dtype=torch.float32, device=self.device)
        mats = []
        for c, ce in enumerate(self.ces):
            vals, inv = torch.unique(ce, return_inverse=True)
            if vals.numel() > 256:
                return
            lut[c, :vals.numel()] = vals
            mats.append(inv.to(torch.uint8).contiguous())
        self.ce_mat, self.ce_lut = mats, lut

    def set_ch(self, chx, chy, chz):
        self.chs = [self._fortran(c) for c in (chx, chy, chz)]

    def set_field(self, name, arr):
        self.ehs[_NAMES.index(name)].copy_(self._fortran(arr))

    def field(self, name):
        """(nx, ny, nz) view (Fortran strides) of a field."""
        return self.ehs[_NAMES.index(name)].permute(2, 1, 0)

    def add_source(self, str_f, pt0, pt1, tfunc, spatial_value=1.0):
        """Soft source on an E component over the inclusive box pt0..pt1, applied after
        every E stage with the value float32(tfunc(tstep + s/5)) * spatial_value."""
        if str_f not in ("ex", "ey", "ez"):
            raise ValueError(f"4th-order sources drive E components : {str_f}")
        p0, p1 = tuple(int(v) for v in pt0), tuple(int(v) for v in pt1)
        n = (self.nx, self.ny, self.nz)
        if any(not (0 <= a <= b < m) for a, b, m in zip(p0, p1, n)):
            raise ValueError(f"source box {p0}..{p1} outside the grid {n}")
        sv = None
        svs = 1.0
        if isinstance(spatial_value, (np.ndarray, torch.Tensor)):
            shp = tuple(b - a + 1 for a, b in zip(p0, p1))
            sv_np = np.asarray(spatial_value, np.float32).reshape(shp)
            # box array in the (k, j, i) order of the memory view
            sv = torch.tensor(np.ascontiguousarray(sv_np.transpose(2, 1, 0)), device=self.device)
        else:
            svs = float(np.float32(spatial_value))
        self.sources.append((_NAMES.index(str_f), p0, p1, tfunc, sv, svs))

    def add_probe(self, str_f, pt):
        i, j, k = (int(v) for v in pt)
        if not (0 <= i < self.nx and 0 <= j < self.ny and 0 <= k < self.nz):
            raise ValueError(f"probe point {pt} outside the grid")
        self.probe_points.append((_NAMES.index(str_f), (i, j, k)))

    # -------------------------------------------------------------- stepping
    def _h(self, s, st):
        e = [t.data_ptr() for t in self.ehs]
        ch = [c.data_ptr() for c in self.chs] if self.chs is not None else [None] * 3
        _lib.call("mifdtd_s4_update_h", self.nx, self.ny, self.nz, self.cl[s], *e, *ch, self.c0, st)
        if self.npml:
            d = self._half(self.psiH, *self._tabH[s])
            _lib.call("mifdtd_s4_cpml_h", self.nx, self.ny, self.nz, ctypes.byref(d), self.cl[s], *e, *ch,
                      self.c0, st)

    def _e(self, s, st, events=None):
        e = [t.data_ptr() for t in self.ehs]
        ce = [c.data_ptr() for c in self.ces] if self.ces is not None else [None] * 3
        if events is not None:
            events[0].record(self.stream)
        if self.ce_mat is not None:
            _lib.call("mifdtd_s4_update_e_lut", self.nx, self.ny, self.nz, self.dl[s], *e,
                      *[m.data_ptr() for m in self.ce_mat], self.ce_lut.data_ptr(), st)
        else:
            _lib.call("mifdtd_s4_update_e", self.nx, self.ny, self.nz, self.dl[s], *e, *ce, self.c0, st)
        if events is not None:
            events[1].record(self.stream)
        if self.npml:
            d = self._half(self.psiE, *self._tabE[s])
            _lib.call("mifdtd_s4_cpml_e", self.nx, self.ny, self.nz, ctypes.byref(d), self.dl[s], *e, *ce,
                      self.c0, st)
        for c, p0, p1, tfunc, sv, svs in self.sources:
            tv = float(np.float32(tfunc(self.tstep + s / 5.0)))
            # memory view (nz, ny, nx): box (k, j, i)
            _lib.call("mifdtd_source", self.ehs[c].data_ptr(), self.nz, self.ny, self.nx,
                      p0[2], p1[2], p0[1], p1[1], p0[0], p1[0], tv,
                      None if sv is None else sv.data_ptr(), svs, 0, st)

    def step(self, buf=None, t=0, events=None):
        """One step; ``events`` (a pair of timing events) brackets the first E-stage
        kernel on the simulation's stream."""
        st = _lib.stream_handle(self.stream)
        for s in range(4):
            self._h(s, st)
            self._e(s, st, events if s == 0 else None)
        self._h(4, st)
        self.tstep += 1
        if buf is not None:
            for n, (c, (i, j, k)) in enumerate(self.probe_points):
                buf[t, n] = self.ehs[c][k, j, i]

    def run(self, nsteps):
        nprobe = len(self.probe_points)
        buf = torch.zeros((nsteps, nprobe), dtype=torch.float32, device=self.device) if nprobe else None
        for t in range(nsteps):
            self.step(buf, t)
        if buf is None:
            return None
        s = buf.cpu().numpy()
        self._series.append(s)
        return s

    def series(self):
        return np.concatenate(self._series) if self._series else None
