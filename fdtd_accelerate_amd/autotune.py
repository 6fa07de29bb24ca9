"""Tile-shape selection for the whole-step kernels.

``slab_planes`` returns the x planes per slab of the slab-ordered step: a table of
measured winners (``tune_slab``, file ``tuned_slab.json``) keyed by
(nx, ny, nz_pitch, npml, ce), else the whole grid as one slab.  (Re-reading a slab's
E and H from the 256 MB last-level cache was measured at only ~10% above HBM
bandwidth on MI355X, less than the per-launch cost of small slabs.)

``fused_tiles`` returns (rows_per_thread T, planes_per_chunk CX) for a grid: a table of
measured winners (filled by ``tune_fused``) keyed by (nx, ny, nz_pitch, npml), else a
heuristic that keeps enough work-groups in flight (>= 4 per CU on 256 CUs) while
keeping the warm-up plane overhead (1/CX) small.
"""
from __future__ import annotations

import json
import os

_TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_tiles.json")
_table = None
N_CU = 256


def _load():
    global _table
    if _table is None:
        try:
            with open(_TABLE_PATH) as fh:
                _table = {tuple(json.loads(k)): tuple(v) for k, v in json.load(fh).items()}
        except (OSError, ValueError):
            _table = {}
    return _table


def lds_fits(T, CX, table, ntri=0, extra=0):
    """Whether the sweep's LDS (E and H row exchange 2 x 2 x T x 2 x 256 floats, the k-1
    carry (CX+1)(T+1)2 floats, the 3 x 256 CE table or ``ntri`` CE triples, ``extra``
    more floats, ~0.4 KiB of source/probe tables) fits the 160 KiB of one gfx950
    work-group."""
    tab = 4 * ((3 * ntri + 3) // 4) if ntri else (768 if table else 0)
    return 4 * (tab + 2 * 2 * T * 2 * 256 + (CX + 1) * (T + 1) * 2 + extra) + 512 <= 160 * 1024


def lds_fits_pairs(T, CX, table):
    """Whether the two-step sweep's LDS (T+3 rows: H and E row exchanges of both levels,
    4 x (T+3) x 2 x 256 floats; the k-1 carries of both levels per plane of the chunk,
    (CX+3) x (T+3) x 4 floats; the CE table; ~0.7 KiB of source/probe tables) fits the
    160 KiB of one gfx950 work-group."""
    R = T + 3
    return 4 * ((768 if table else 0) + 4 * R * 512 + 4 * (-(-R // 4)) + (CX + 3) * R * 4) + 1024 <= 160 * 1024


def heuristic(nx, ny, nzp, npml):
    """T owned rows per block (y halo overhead 1/T; the block has T+1 waves), CX planes
    per chunk (seed-plane overhead 1/CX), keeping >= 1024 blocks.  T = 9 measured best
    at 1024^3 (10.87 ms per sweep against 11.40 for T = 7 and 10.98 for T = 11, CX = 64,
    interleaved on one MI355X): a smaller halo share outweighs the lower occupancy of
    one 10-wave work-group per CU."""
    T = min(9, max(1, ny // 8))
    nrb = -(-ny // T)
    cx = min(nx, 64)
    while cx > 8 and nrb * (-(-nx // cx)) < 1024:
        cx = -(-cx // 2)
    return T, max(1, cx)


def heuristic_pairs(nx, ny, nzp, npml):
    """(T, CX) of the two-step sweep: T = 9 owned rows (12 waves, the kernel's launch
    bound; 9.95 ms per step at 1024^3 against 11.8 for T = 5), CX = 64 planes, halved
    while fewer than 1024 blocks."""
    T = min(9, max(1, ny // 8))
    nrb = -(-ny // T)
    cx = min(nx, 64)
    while cx > 8 and nrb * (-(-nx // cx)) < 1024:
        cx = -(-cx // 2)
    while cx > 1 and not lds_fits_pairs(T, cx, True):
        cx -= 1
    return T, max(1, cx)


def fused_pairs(nx, ny, nzp, npml):
    """Whether the fused engine advances steps in pairs by default on this grid: never
    (at 1024^3 with CPML 12 a pair takes 27.3 ms against 25.1 ms for two single steps on
    MI355X; the two-step sweep halves the bytes but is issue-bound)."""
    return False


def fused_tiles(nx, ny, nzp, npml, pairs=False):
    if pairs:
        return heuristic_pairs(nx, ny, nzp, npml)
    t = _load().get((nx, ny, nzp, npml))
    return t if t is not None else heuristic(nx, ny, nzp, npml)


def tune_fused(sim, candidates=((7, 64), (9, 64), (9, 96), (11, 64), (11, 128), (13, 64), (15, 96)), steps=5, save=True):
    """Time the interior sweep for each (T, CX) on ``sim`` (a YeeSim with engine='fused')
    and record the best.  Runs real steps (the simulation advances by
    len(candidates) * (steps + 1) steps)."""
    from .fused import FusedStepper
    f = sim.fields
    best = None
    results = {}
    for T, CX in candidates:
        if not lds_fits(T, CX, f.ce_mat is not None):
            continue
        st = FusedStepper(sim, rows_per_thread=T, planes_per_chunk=CX)
        st.run(1, None)
        ms = st.run(steps, None, time_kernel=True)
        results[(T, CX)] = ms
        if best is None or ms < results[best]:
            best = (T, CX)
    if best is None:
        raise ValueError(f"no candidate tile fits the 160 KiB of LDS : {tuple(candidates)}")
    key = (f.nx, f.ny, f.nz_pitch, sim.cpml.npml if sim.cpml else 0)
    _load()[key] = best
    if save:
        with open(_TABLE_PATH, "w") as fh:
            json.dump({json.dumps(list(k)): list(v) for k, v in _table.items()}, fh, indent=1)
    return best, results


_SLAB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_slab.json")
_slab_table = None


def _load_slab():
    global _slab_table
    if _slab_table is None:
        try:
            with open(_SLAB_PATH) as fh:
                _slab_table = {tuple(json.loads(k)): int(v) for k, v in json.load(fh).items()}
        except (OSError, ValueError):
            _slab_table = {}
    return _slab_table


def slab_heuristic(nx, ny, nzp):
    return nx


def slab_planes(nx, ny, nzp, npml, ce_on):
    t = _load_slab().get((nx, ny, nzp, npml, int(bool(ce_on))))
    return t if t is not None else slab_heuristic(nx, ny, nzp)


def tune_slab(sim, candidates=(1, 2, 3, 4, 6, 8, 12, 16, 32), steps=5, save=True):
    """Time each slab size on ``sim`` (advances the simulation) and record the best."""
    from .fused import SlabStepper
    f = sim.fields
    results = {}
    for S in candidates:
        if S > f.nx:
            continue
        st = SlabStepper(sim, slab_planes=S)
        st.run(1, None)
        results[S] = st.run(steps, None, time_kernel=True)
    if not results:
        raise ValueError(f"no candidate slab depth is at most nx = {f.nx} : {tuple(candidates)}")
    best = min(results, key=results.get)
    key = (f.nx, f.ny, f.nz_pitch, sim.cpml.npml if sim.cpml else 0, int(f.ce_on))
    _load_slab()[key] = best
    if save:
        with open(_SLAB_PATH, "w") as fh:
            json.dump({json.dumps(list(k)): v for k, v in _slab_table.items()}, fh, indent=1)
    return best, results
