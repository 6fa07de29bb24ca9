"""Writes tests/golden/adjoint_design160.npz: FP64 reference data for the adjoint gradient
of the 160^3 design problem (run here, on the CPU, with the oracle):

  160^3 grid, CPML 10, 64^3 design box at the centre with eps_r ~ U[1,4] (float32,
  numpy.random.default_rng(EPS_SEED)), Ez soft line source on the plane x = 20 (the line
  y = 80 along z over the non-CPML interior) driven by the Gaussian-modulated pulse, objective J = sum over steps 601..800
  of sum over the plane x = 140 of |E|^2.

* central finite differences of the float64 C oracle (OracleSim) at 32 seeded voxels,
  with the step h chosen by a sweep on three voxels (the h whose estimate changes least
  against h/3);
* the float64 gradient on the whole box from torch.autograd through the float64
  restatement oracle/torch_ref.py (checkpointed).

    python tests/golden/make_adjoint_golden.py [--quick] [--device cuda]

The finite differences run on the CPU (the C oracle) and are kept in <out>.fd.npz until
the script ends, so a rerun (e.g. with --device cuda for the autograd stage, whose
float64 graph of the 160^3 problem needs more host memory than a small machine has)
skips them.
"""
import argparse
import os
import sys
import time

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from oracle.torch_ref import TorchYee, objective_grad  # noqa: E402

EPS_SEED, VOX_SEED = 4, 44


def config(quick):
    if quick:
        return dict(n=48, npml=6, box=(16, 16, 16, 16, 16, 16), src=8, mon=38, nsteps=200, window=60, nvox=6)
    return dict(n=160, npml=10, box=(48, 48, 48, 64, 64, 64), src=20, mon=140, nsteps=800, window=200, nvox=32)


def design_eps(box):
    return np.random.default_rng(EPS_SEED).uniform(1.0, 4.0, box[3:]).astype(np.float32)


def voxels(box, count):
    r = np.random.default_rng(VOX_SEED)
    return np.stack([r.integers(0, box[3 + a], count) for a in range(3)], 1)


def oracle_J(cfg, eps64):
    n, npml, box = cfg["n"], cfg["npml"], cfg["box"]
    i0, j0, k0, di, dj, dk = box
    ces = [np.full((n, n, n), 0.5) for _ in range(3)]
    for c in ces:
        c[i0:i0 + di, j0:j0 + dj, k0:k0 + dk] = 0.5 / eps64
    o = O.OracleSim(n, n, n, 0, npml, np.float64, ces=ces)
    o.add_source("ez", (cfg["src"], n // 2, npml), (cfg["src"], n // 2, n - npml - 1), O.pulse)
    n0 = cfg["nsteps"] - cfg["window"] + 1
    J = 0.0
    for s in range(1, cfg["nsteps"] + 1):
        o.step()
        if s >= n0:
            m = cfg["mon"]
            J += float((o.ehs[0][m] ** 2).sum() + (o.ehs[1][m] ** 2).sum() + (o.ehs[2][m] ** 2).sum())
    return J


def fd(cfg, eps64, v, h):
    e1 = eps64.copy()
    e1[tuple(v)] += h
    e2 = eps64.copy()
    e2[tuple(v)] -= h
    return (oracle_J(cfg, e1) - oracle_J(cfg, e2)) / (2 * h)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--out", default=os.path.join(HERE, "adjoint_design160.npz"))
    ap.add_argument("--device", default="cpu", help="where the float64 autograd runs (cpu or cuda)")
    ap.add_argument("--fd-only", action="store_true", help="stop after the finite-difference stage")
    a = ap.parse_args()
    cfg = config(a.quick)
    torch.set_num_threads(os.cpu_count() or 8)
    eps32 = design_eps(cfg["box"])
    eps64 = eps32.astype(np.float64)
    vox = voxels(cfg["box"], cfg["nvox"])
    stage = a.out + ".fd.npz"  # finite-difference stage, kept so a rerun skips it
    if os.path.exists(stage):
        z = np.load(stage)
        J, h, hs, sweep, fds = float(z["J64"]), float(z["h"]), list(z["hs"]), z["sweep"], z["fd"]
        print("finite differences loaded from", stage, flush=True)
    else:
        t0 = time.time()
        J = oracle_J(cfg, eps64)
        print(f"J64 {J:.12g}  ({time.time() - t0:.1f} s per run)", flush=True)
        hs = [3e-2, 1e-2, 3e-3, 1e-3, 3e-4, 1e-4, 3e-5]
        sweep = np.zeros((3, len(hs)))
        for a_, v in enumerate(vox[:3]):
            for b_, h in enumerate(hs):
                sweep[a_, b_] = fd(cfg, eps64, v, h)
            print("sweep", v, sweep[a_], flush=True)
        change = np.abs(np.diff(sweep, axis=1)) / np.abs(sweep[:, 1:])
        h = hs[int(np.argmin(change.max(0))) + 1]
        print("h", h, flush=True)
        fds = np.array([fd(cfg, eps64, v, h) for v in vox])
        np.savez(stage, J64=J, h=h, hs=np.array(hs), sweep=sweep, fd=fds)
        print("fd done", flush=True)
    if a.fd_only:
        return
    n, npml, box = cfg["n"], cfg["npml"], cfg["box"]
    dev = torch.device(a.device)
    m = TorchYee(n, n, n, 0, npml, device=dev)
    ceb = [torch.full((n, n, n), 0.5, dtype=torch.float64, device=dev) for _ in range(3)]
    mask = torch.zeros((n, n, n), dtype=torch.float64, device=dev)
    mask[cfg["src"], n // 2, npml:n - npml] = 1
    n0 = cfg["nsteps"] - cfg["window"] + 1
    Jt, g = objective_grad(m, ceb, box, torch.tensor(eps64, device=dev), cfg["nsteps"], n0, cfg["mon"],
                           lambda s: [(2, O.pulse(s) * mask)], seg=2, outer=None if a.quick else 50,
                           progress=lambda msg: print(msg, flush=True))
    print(f"autograd J {Jt:.12g} (oracle {J:.12g})", flush=True)
    g = g.cpu().numpy()
    rel = np.abs(g[tuple(vox.T)] - fds) / np.abs(fds)
    print("autograd vs fd max rel", rel.max(), flush=True)
    np.savez_compressed(a.out, config=np.array(repr(cfg)), eps_seed=EPS_SEED, vox_seed=VOX_SEED,
                        eps_sum=float(eps64.sum()), voxels=vox, h=h, hs=np.array(hs), sweep=sweep, fd=fds,
                        J64=J, grad_ref=g.astype(np.float32))
    print("wrote", a.out, flush=True)
    if os.path.exists(stage):
        os.remove(stage)


if __name__ == "__main__":
    main()
