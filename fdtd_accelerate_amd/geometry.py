"""Synthetic media of the benchmark configurations: seeded non-overlapping dielectric
spheres, sampled onto the CE arrays at each E component's Yee position.

Positions on the KEMP grid (the forward-difference staggering of K7/cpu/src/core.c,
equivalently ``KU/cpu/matter.py:74-79`` without kuFDTD's ghost offset):
  Ex(i,j,k) at (i, j+1/2, k+1/2),  Ey at (i+1/2, j, k+1/2),  Ez at (i+1/2, j+1/2, k).
CE = 0.5 / eps_r there (``KU/cpu/matter.py:81-83``: dt/dx = 0.5 in normalised units).
The reference has only cylinders (``KU/structure.py:85-139``); spheres are this
package's own generator.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch


@dataclasses.dataclass
class Sphere:
    center: tuple
    radius: float
    eps_r: float


def random_spheres(seed, count, r_range, eps_range, lo, hi, max_tries=200000):
    """``count`` non-overlapping spheres (rejection sampling, ``numpy.random.default_rng(seed)``):
    radius ~ U[r_range], eps_r ~ U[eps_range], centre ~ U so that the whole sphere lies
    in the box [lo, hi] (per-axis cell coordinates).  Raises if the box cannot hold them."""
    rng = np.random.default_rng(seed)
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    out = []
    tries = 0
    while len(out) < count:
        tries += 1
        if tries > max_tries:
            raise RuntimeError(f"could only place {len(out)} of {count} spheres")
        r = rng.uniform(*r_range)
        if np.any(hi - lo < 2 * r):
            continue
        c = rng.uniform(lo + r, hi - r)
        if any(np.sum((c - s.center) ** 2) < (r + s.radius) ** 2 for s in out):
            continue
        out.append(Sphere(tuple(float(x) for x in c), float(r), float(rng.uniform(*eps_range))))
    return out


_OFFSETS = ((0.0, 0.5, 0.5), (0.5, 0.0, 0.5), (0.5, 0.5, 0.0))


def paint_ce(ces, spheres, ns, background=1.0):
    """Write CE = 0.5/eps_r of ``spheres`` (on ``background``) into the three CE tensors
    ``(nx, ny, nz_pitch)`` (device or host torch tensors); only each sphere's bounding box
    is touched after the background fill."""
    nx, ny, nz = ns
    for c, ce in enumerate(ces):
        ce.fill_(0.5 / background)
        off = _OFFSETS[c]
        for s in spheres:
            lo = [max(0, int(np.floor(s.center[a] - s.radius - off[a]))) for a in range(3)]
            hi = [min(n, int(np.ceil(s.center[a] + s.radius - off[a])) + 1) for a, n in enumerate((nx, ny, nz))]
            if any(h <= l for l, h in zip(lo, hi)):
                continue
            dev = ce.device
            ax = [torch.arange(lo[a], hi[a], device=dev, dtype=torch.float64) + off[a] - s.center[a] for a in range(3)]
            d2 = ax[0][:, None, None] ** 2 + ax[1][None, :, None] ** 2 + ax[2][None, None, :] ** 2
            box = ce[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
            box[d2 <= s.radius ** 2] = float(np.float32(0.5 / s.eps_r))


def paint_ce_numpy(ns, nz_pitch, spheres, background=1.0):
    """Host (numpy) CE arrays of the same media, for CPU checks."""
    out = [torch.zeros((ns[0], ns[1], nz_pitch), dtype=torch.float32) for _ in range(3)]
    paint_ce(out, spheres, ns, background)
    return [t.numpy() for t in out]


def eps_map_numpy(ns, spheres, background=1.0):
    """eps_r sampled at the integer points (i, j, k) of an (nx, ny, nz) grid: the
    single per-cell permittivity map of the 4th-order path (one coefficient array per
    field, ``S4/PyCUDA/100-base.py:75,84-89``)."""
    nx, ny, nz = ns
    eps = np.full((nx, ny, nz), background, np.float32)
    for s in spheres:
        lo = [max(0, int(np.floor(s.center[a] - s.radius))) for a in range(3)]
        hi = [min(n, int(np.ceil(s.center[a] + s.radius)) + 1) for a, n in enumerate(ns)]
        if any(h <= l for l, h in zip(lo, hi)):
            continue
        ax = [np.arange(lo[a], hi[a], dtype=np.float64) - s.center[a] for a in range(3)]
        d2 = ax[0][:, None, None] ** 2 + ax[1][None, :, None] ** 2 + ax[2][None, None, :] ** 2
        box = eps[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
        box[d2 <= s.radius ** 2] = np.float32(s.eps_r)
    return eps
