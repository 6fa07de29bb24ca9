"""ORACLE — test infrastructure only (imported by tests/ and tests/golden scripts).

Float64 CPU restatement of the in-place Yee + CPML step written with out-of-place torch
operations, so that torch.autograd gives the exact gradient of the FP64 model: the
reference gradient the hand-written adjoint kernels are checked against.

Semantics follow the C oracle (``oracle_body.h``): core update of
``K7/cpu/src/core.c:11-105`` with its masks (walls untouched, masked lanes zeroed),
CPML faces of ``KU/cpu/core/cpml_non_kappa.c`` (x, y, z order; kuFDTD tables; applied
where the core writes), soft sources after the E half (``K7/cpu/incident_direct.py``),
CH = 0.5.  The masked lanes read a zero instead of the next row's first element: the
value is multiplied by 0 there, so the result (and the gradient) is the same.
"""
from __future__ import annotations

import numpy as np
import torch
from torch.utils.checkpoint import checkpoint

from oracle import oracle as O

DT = torch.float64


def _fwd(f, dim):
    """f[+1] - f along dim, zero in the last slice."""
    d = f.narrow(dim, 1, f.shape[dim] - 1) - f.narrow(dim, 0, f.shape[dim] - 1)
    pad = torch.zeros_like(f.narrow(dim, 0, 1))
    return torch.cat([d, pad], dim)


def _bwd(f, dim):
    """f - f[-1] along dim, zero in the first slice."""
    d = f.narrow(dim, 1, f.shape[dim] - 1) - f.narrow(dim, 0, f.shape[dim] - 1)
    pad = torch.zeros_like(f.narrow(dim, 0, 1))
    return torch.cat([pad, d], dim)


class TorchYee:
    """(nx, ny, nzp) float64 fields, pad lanes = nzp - nz, CPML npml on all faces."""

    def __init__(self, nx, ny, nz, pad=0, npml=0, R=1e-6, m=4, device="cpu"):
        self.nx, self.ny, self.nz, self.pad, self.npml = nx, ny, nz, pad, npml
        nzp = nz + pad
        self.nzp = nzp
        self.device = torch.device(device)
        dev = self.device
        I = torch.arange(nx, device=dev).view(-1, 1, 1)
        J = torch.arange(ny, device=dev).view(1, -1, 1)
        K = torch.arange(nzp, device=dev).view(1, 1, -1)
        full = torch.ones((nx, ny, nzp), dtype=torch.bool, device=dev)
        # rows written by the core and lanes kept (valid) inside those rows
        self.erow = [(J < ny - 1) & full, (I < nx - 1) & full, (I < nx - 1) & (J < ny - 1) & full]
        self.elane = [(K < nz - 1) & full, (K < nz - 1) & full, (K < nz) & full]
        self.hrow = [(J > 0) & full, (I > 0) & full, (I > 0) & (J > 0) & full]
        self.hlane = [(K > 0) & (K < nz) & full, (K > 0) & (K < nz) & full, (K < nz) & full]
        self.evalid = [r & l for r, l in zip(self.erow, self.elane)]
        self.hvalid = [r & l for r, l in zip(self.hrow, self.hlane)]
        if npml:
            bE, aE, bH, aH = O.cpml_tables(npml, m, R, dt=0.5, dtype=np.float64)
            self.tab = {k: torch.tensor(v, dtype=DT, device=dev) for k, v in (("bE", bE), ("aE", aE), ("bH", bH), ("aH", aH))}

    def zeros(self):
        z = lambda: torch.zeros((self.nx, self.ny, self.nzp), dtype=DT, device=self.device)
        st = [z() for _ in range(6)]
        if self.npml:
            st += [torch.zeros(self._psi_shape(ax), dtype=DT, device=self.device) for ax in range(3) for _ in range(2)
                   for _ in range(2)]
            st += [torch.zeros(self._psi_shape(ax), dtype=DT, device=self.device) for ax in range(3) for _ in range(2)
                   for _ in range(2)]
        return st

    def _psi_shape(self, ax):
        return [(self.npml, self.ny, self.nzp), (self.nx, self.npml, self.nzp), (self.nx, self.ny, self.npml)][ax]

    def _slab(self, ax, side, half):
        """indices along axis ax of the slab planes and their table entries"""
        n = (self.nx, self.ny, self.nz)[ax]
        p = np.arange(self.npml)
        if half == "E":
            q = n - 1 - self.npml + p if side else p
            t = self.npml + 1 + p if side else p + 1
        else:
            q = n - self.npml + p if side else p
            t = self.npml + 2 + p if side else p + 1
        return torch.tensor(q, device=self.device), torch.tensor(t, device=self.device)

    @staticmethod
    def _bc(v, ax):
        shp = [1, 1, 1]
        shp[ax] = -1
        return v.view(shp)

    def step(self, st, ce, src):
        """One step.  st: list of state tensors (6 fields [+ 24 psi]); ce: 3 CE tensors;
        src: list of (component, additive tensor of full shape) for this step."""
        ex, ey, ez, hx, hy, hz = st[:6]
        psi = st[6:]
        # ---------------------------------------------------------------- E core
        cur = [_fwd(hz, 1) - _fwd(hy, 2), _fwd(hx, 2) - _fwd(hz, 0), _fwd(hy, 0) - _fwd(hx, 1)]
        E = [ex, ey, ez]
        E = [torch.where(self.evalid[c], E[c] + ce[c] * cur[c], torch.where(self.erow[c], torch.zeros_like(E[c]), E[c]))
             for c in range(3)]
        H = [hx, hy, hz]
        newpsi = list(psi)
        if self.npml:
            # (axis -> (E comp of psi1, H driving it, sign), (E comp of psi2, H, sign))
            pairs = {0: ((1, 2, -1.0), (2, 1, +1.0)), 1: ((2, 0, -1.0), (0, 2, +1.0)), 2: ((0, 1, -1.0), (1, 0, +1.0))}
            for ax in range(3):
                for sd in range(2):
                    q, t = self._slab(ax, sd, "E")
                    b = self._bc(self.tab["bE"][t], ax)
                    a = self._bc(self.tab["aE"][t], ax)
                    for pc, (ec, hc, sg) in enumerate(pairs[ax]):
                        k = (ax * 2 + sd) * 2 + pc
                        d = _fwd(H[hc], ax).index_select(ax, q)
                        v = self.evalid[ec].index_select(ax, q)
                        new = torch.where(v, b * psi[k] + a * d, psi[k])
                        newpsi[k] = new
                        upd = torch.where(v, ce[ec].index_select(ax, q) * new, torch.zeros_like(new))
                        E[ec] = E[ec].index_add(ax, q, sg * upd)
        for c, add in src:
            E[c] = E[c] + add
        # ---------------------------------------------------------------- H core
        cur = [_bwd(E[2], 1) - _bwd(E[1], 2), _bwd(E[0], 2) - _bwd(E[2], 0), _bwd(E[1], 0) - _bwd(E[0], 1)]
        H = [torch.where(self.hvalid[c], H[c] - 0.5 * cur[c], torch.where(self.hrow[c], torch.zeros_like(H[c]), H[c]))
             for c in range(3)]
        if self.npml:
            pairs = {0: ((1, 2, +1.0), (2, 1, -1.0)), 1: ((2, 0, +1.0), (0, 2, -1.0)), 2: ((0, 1, +1.0), (1, 0, -1.0))}
            for ax in range(3):
                for sd in range(2):
                    q, t = self._slab(ax, sd, "H")
                    b = self._bc(self.tab["bH"][t], ax)
                    a = self._bc(self.tab["aH"][t], ax)
                    for pc, (hc, ec, sg) in enumerate(pairs[ax]):
                        k = 12 + (ax * 2 + sd) * 2 + pc
                        d = _bwd(E[ec], ax).index_select(ax, q)
                        v = self.hvalid[hc].index_select(ax, q)
                        new = torch.where(v, b * psi[k] + a * d, psi[k])
                        newpsi[k] = new
                        upd = torch.where(v, 0.5 * new, torch.zeros_like(new))
                        H[hc] = H[hc].index_add(ax, q, sg * upd)
        return E + H + newpsi


def objective_grad(model, ce_background, box, eps, nsteps, n0, monitor, sources, seg=10, outer=None,
                   want_grad=True, progress=None):
    """J and dJ/deps (float64) for the design box ``box`` = (i0, j0, k0, di, dj, dk).

    ce_background: 3 float64 CE tensors; on the box, CE = 0.5/eps on all components.
    sources: function(step n) -> list of (component, additive tensor).
    Memory: without ``outer`` the steps run in checkpointed segments of ``seg`` steps
    (one saved state per segment).  With ``outer`` the forward keeps a state every
    ``outer`` steps only and the reverse sweep recomputes each outer segment (with
    ``seg``-step checkpoints inside), carrying the adjoint of the state between
    segments: the gradient is the same, the peak memory ~ (outer/seg + 1) states.
    ``progress``: optional function called with a message per outer segment."""
    i0, j0, k0, di, dj, dk = box
    eps = torch.as_tensor(eps, dtype=DT).clone().requires_grad_(want_grad)

    def make_ce():
        ce = []
        for c in range(3):
            base = ce_background[c]
            padding = (k0, base.shape[2] - k0 - dk, j0, base.shape[1] - j0 - dj, i0, base.shape[0] - i0 - di)
            inner = torch.nn.functional.pad(0.5 / eps, padding)
            mask = torch.nn.functional.pad(torch.ones((di, dj, dk), dtype=DT, device=eps.device), padding)
            ce.append(base * (1 - mask) + inner)
        return ce

    def run_seg(ce, n_start, n_end, *state):
        st = list(state)
        acc = torch.zeros((), dtype=DT, device=eps.device)
        for n in range(n_start, n_end):
            st = model.step(st, ce, sources(n))
            if n >= n0:
                acc = acc + (st[0][monitor] ** 2).sum() + (st[1][monitor] ** 2).sum() + (st[2][monitor] ** 2).sum()
        return (acc, *st)

    def chunked(ce, s0, s1, st, grad):
        J = torch.zeros((), dtype=DT, device=eps.device)
        for s in range(s0, s1, seg):
            e = min(s + seg, s1)
            if grad:
                out = checkpoint(run_seg, ce, s, e, *st, use_reentrant=False)
            else:
                out = run_seg(ce, s, e, *st)
            J = J + out[0]
            st = list(out[1:])
        return J, st

    if not want_grad:
        with torch.no_grad():
            J, _ = chunked(make_ce(), 1, nsteps + 1, model.zeros(), False)
        return float(J), None
    if outer is None:
        J, _ = chunked(make_ce(), 1, nsteps + 1, model.zeros(), True)
        J.backward()
        return float(J.detach()), eps.grad.detach().clone()
    # two levels: states every `outer` steps, reverse sweep over the outer segments
    bounds = list(range(1, nsteps + 1, outer)) + [nsteps + 1]
    saved = {}
    J = 0.0
    with torch.no_grad():
        ce0 = make_ce()
        st = model.zeros()
        for a, b in zip(bounds[:-1], bounds[1:]):
            saved[a] = [t.clone() for t in st]
            Jp, st = chunked(ce0, a, b, st, False)
            J += float(Jp)
        del ce0, st
    lam = None
    for a, b in reversed(list(zip(bounds[:-1], bounds[1:]))):
        if progress:
            progress(f"reverse sweep: steps {a}..{b - 1}")
        st0 = [t.clone().requires_grad_(True) for t in saved.pop(a)]
        ce = make_ce()
        Jp, st1 = chunked(ce, a, b, st0, True)
        outs, grads = [], []
        if Jp.requires_grad:  # segments before the objective window add nothing to J
            outs.append(Jp)
            grads.append(torch.ones((), dtype=DT, device=eps.device))
        if lam is not None:
            outs += st1
            grads += lam
        if outs:
            torch.autograd.backward(outs, grads)
        lam = [t.grad if t.grad is not None else torch.zeros_like(t) for t in st0]
        del st0, st1, ce, outs
    return J, eps.grad.detach().clone()
