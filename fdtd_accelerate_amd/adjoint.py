"""Adjoint gradients of a field objective with respect to the permittivity of a design box.

Problem (``DesignProblem``): a Yee + CPML simulation on the in-place engine's
arithmetic (here stepped by the slab engine, which is bit-identical) with CE = 0.5 in
vacuum and CE = 0.5/eps_r on every E component of each design voxel, soft sources,
and the objective

    J = sum over steps n in [n0, nsteps] of sum over the monitor plane x = im of
        ex^2 + ey^2 + ez^2   (E after step n)

``forward(eps)`` runs the steps once, keeping a checkpoint of the full state (fields
and psi) every ``ckpt`` steps and the monitor-plane E of the objective window.
``backward()`` walks the segments in reverse: it restores a checkpoint, recomputes the
segment (storing the curl of H on the design box before each step) and runs the
hand-written adjoint step (``mifdtd_adjoint_step``) back through it, accumulating
dJ/dCE in double precision; dJ/deps = -0.5/eps^2 * sum over the three components.
Memory is bounded by one segment of curls plus the checkpoints.  ``FdtdObjective`` is
the torch.autograd.Function around it.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .fused import SlabStepper
from .sim import YeeSim
from .sources import gaussian_pulse


class DesignProblem:
    def __init__(self, n=(160, 160, 160), npml=10, box=(48, 48, 48, 64, 64, 64), src_plane=20,
                 monitor=140, nsteps=800, window=200, ckpt=50, tfunc=None, src_box=None):
        nx, ny, nz = n
        self.n = n
        self.npml = npml
        self.box = tuple(int(v) for v in box)
        self.monitor = int(monitor)
        self.nsteps = int(nsteps)
        self.n0 = self.nsteps - int(window) + 1  # first step of the objective window
        self.ckpt = int(ckpt)
        self.tfunc = tfunc if tfunc is not None else gaussian_pulse()
        i0, j0, k0, di, dj, dk = self.box
        if not (0 < i0 and i0 + di < nx and 0 < j0 and j0 + dj < ny - 1 and 0 < k0 and k0 + dk < nz - 1):
            raise ValueError(f"design box {self.box} must lie inside the grid {n}")
        # the adjoint's gradient has no term for the CPML correction of E (-/+CE*psi), so
        # the box keeps off every E slab plane: [npml, n-1-npml) on each axis
        if npml and not (i0 >= npml and i0 + di <= nx - 1 - npml and j0 >= npml and j0 + dj <= ny - 1 - npml
                         and k0 >= npml and k0 + dk <= nz - 1 - npml):
            raise ValueError(f"design box {self.box} overlaps a CPML slab (npml {npml}) of the grid {n}")
        self.sim = YeeSim(nx, ny, nz, npml=npml, coeff_use="e")
        # Ez soft line source on the plane x = src_plane: the line y = ny/2 along z over the
        # non-CPML interior
        if src_box is None:
            src_box = ((src_plane, ny // 2, npml), (src_plane, ny // 2, nz - npml - 1))
        self.src = self.sim.add_source("ez", src_box[0], src_box[1], self.tfunc)
        self.stepper = SlabStepper(self.sim)
        f = self.sim.fields
        self.dbox = _lib.DesignBox(*self.box)
        self.dev = f.device
        self.stream = f.stream
        self._eps = None

    # -------------------------------------------------------------- state helpers
    def _state(self):
        f = self.sim.fields
        cp = self.sim.cpml
        return list(f.ehs) + ([cp.psiE_buf, cp.psiH_buf] if cp is not None else [])

    def _reset(self):
        for t in self._state():
            t.zero_()
        self.src.tstep = 1

    def _set_eps(self, eps):
        eps = torch.as_tensor(eps, dtype=torch.float32, device=self.dev)
        i0, j0, k0, di, dj, dk = self.box
        if tuple(eps.shape) != (di, dj, dk):
            raise ValueError(f"eps shape {tuple(eps.shape)} != design box {(di, dj, dk)}")
        ce = 0.5 / eps
        f = self.sim.fields
        for c in f.ces:
            c[i0:i0 + di, j0:j0 + dj, k0:k0 + dk] = ce
        f.ce_mat = f.ce_lut = None  # material indices of the previous CE no longer hold
        self._eps = eps.clone()

    def _step(self):
        self.stepper._call(1, None)

    def _plane(self):
        f = self.sim.fields
        return torch.stack([f.ehs[c][self.monitor] for c in range(3)])  # (3, ny, nzp)

    # ------------------------------------------------------------------- forward
    def forward(self, eps):
        """Run the simulation for eps (design box values); returns J (float)."""
        self._set_eps(eps)
        self._reset()
        self.ckpts = {}
        nwin = self.nsteps - self.n0 + 1
        f = self.sim.fields
        self.planes = torch.empty((nwin, 3, f.ny, f.nz_pitch), dtype=torch.float32, device=self.dev)
        for n in range(1, self.nsteps + 1):
            if (n - 1) % self.ckpt == 0:
                self.ckpts[n - 1] = [t.clone() for t in self._state()]
            self._step()
            if n >= self.n0:
                self.planes[n - self.n0] = self._plane()
        p = self.planes.double()
        return float((p * p).sum())

    # ------------------------------------------------------------------ backward
    def backward(self):
        """dJ/deps on the design box (float64 tensor, box shape) of the last forward."""
        f = self.sim.fields
        cp = self.sim.cpml
        i0, j0, k0, di, dj, dk = self.box
        nbox = di * dj * dk
        lam = [torch.zeros_like(t) for t in f.ehs]
        lpE = torch.zeros_like(cp.psiE_buf) if cp is not None else None
        lpH = torch.zeros_like(cp.psiH_buf) if cp is not None else None
        grad = torch.zeros((3, di, dj, dk), dtype=torch.float64, device=self.dev)
        curls = torch.empty((self.ckpt, 3, di, dj, dk), dtype=torch.float32, device=self.dev)
        lamp = (ctypes.c_void_p * 6)(*[t.data_ptr() for t in lam])
        cep = (ctypes.c_void_p * 3)(*[c.data_ptr() for c in f.ces])
        tabs = (cp.bE.data_ptr(), cp.aE.data_ptr(), cp.bH.data_ptr(), cp.aH.data_ptr()) if cp is not None else (None,) * 4
        st = _lib.stream_handle(self.stream)
        starts = sorted(self.ckpts)
        for s0 in reversed(starts):
            s1 = min(s0 + self.ckpt, self.nsteps)
            for t, c in zip(self._state(), self.ckpts[s0]):
                t.copy_(c)
            self.src.tstep = s0 + 1
            for n in range(s0 + 1, s1 + 1):  # recompute, keeping curl(H^{n-1}) on the box
                _lib.call("mifdtd_design_curl", f.nx, f.ny, f.nz_pitch, f.pad, f.ehs[3].data_ptr(),
                          f.ehs[4].data_ptr(), f.ehs[5].data_ptr(), ctypes.byref(self.dbox),
                          curls[n - s0 - 1].data_ptr(), st)
                self._step()
            for n in range(s1, s0, -1):
                ep = self.planes[n - self.n0].data_ptr() if n >= self.n0 else None
                _lib.call("mifdtd_adjoint_step", f.nx, f.ny, f.nz_pitch, f.pad, ctypes.byref(lamp), ctypes.byref(cep),
                          cp.npml if cp is not None else 0,
                          None if lpE is None else lpE.data_ptr(), None if lpH is None else lpH.data_ptr(),
                          *tabs, ep, self.monitor, ctypes.byref(self.dbox), curls[n - s0 - 1].data_ptr(),
                          grad.data_ptr(), st)
        eps = self._eps.double()
        self.grad_ce = grad
        return grad.sum(0) * (-0.5 / (eps * eps))


class FdtdObjective(torch.autograd.Function):
    """J(eps) with the adjoint gradient: ``FdtdObjective.apply(eps, problem)``."""

    @staticmethod
    def forward(ctx, eps, problem):
        J = problem.forward(eps.detach())
        ctx.problem = problem
        return torch.tensor(J, dtype=torch.float64, device=eps.device)

    @staticmethod
    def backward(ctx, gJ):
        g = ctx.problem.backward()
        return (gJ * g).to(torch.float32), None
