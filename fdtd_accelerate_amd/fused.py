"""Host drivers of the whole-step engines: the fused single-pass step
(``mifdtd_fused_run``) and the slab-ordered step (``mifdtd_slab_run``).

The fused kernel reads the state of step n from one buffer set and writes step n+1 to
a second (ping-pong).  ``FusedStepper`` owns the second field set and the second psi-E
buffer, builds the descriptor from the ``YeeSim`` objects (fields, CPML, sources,
probes), runs K steps in one C call (the C loop launches K kernels back to back) and,
after an odd K, swaps the buffers so that ``sim.fields`` / ``sim.cpml`` always hold
the current state.  Sources advance their tstep exactly as the in-place path does.

The slab step updates in place (no second buffer set): ``SlabStepper`` shares the
descriptor and time-value logic and runs E and H passes slab by slab.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from . import fields as _fields
from .autotune import fused_pairs, fused_tiles, lds_fits, lds_fits_pairs, slab_planes

_COMP = {"ex": 0, "ey": 1, "ez": 2, "hx": 3, "hy": 4, "hz": 5}


def in_place_only():
    """Instance types the in-place engine applies through the Fields instance list; the
    whole-step and batched engines take their step from a descriptor and reject them,
    except those an engine carries itself (the fused engine: one Drude box, one TFSF plane
    wave and periodic boundaries)."""
    from .drude import Drude
    from .pbc import Pbc
    from .tfsf import Tfsf
    return ((Pbc, "periodic boundaries"), (Tfsf, "TFSF source"), (Drude, "Drude medium"))


def _check_whole_step(sim, engine, allow=(), cpml_axes=False):
    """Limits shared by the whole-step engines (same descriptor); ``allow``: instance
    types of ``in_place_only`` this engine carries itself; ``cpml_axes``: CPML may cover
    a subset of the axes (both faces of each or none) instead of all six faces."""
    f = sim.fields
    if f.ch_on:
        raise ValueError(f"the {engine} engine uses CH = 0.5; use engine='kemp' for CH arrays")
    cp = sim.cpml
    if cp is not None:
        both = [set(d) == {"f", "b"} for d in cp.apply_direction]
        if cpml_axes and any(not b and d for b, d in zip(both, cp.apply_direction)):
            raise ValueError(f"the {engine} engine needs CPML on both faces of an axis or none; use engine='kemp'")
        if not cpml_axes and not all(both):
            raise ValueError(f"the {engine} engine needs CPML on all six faces; use engine='kemp'")
    for cls, what in in_place_only():
        if cls not in allow and any(isinstance(i, cls) for i in f.instance_list):
            raise ValueError(f"the {engine} engine has no {what}; use engine='kemp'")
    if len(sim.sources) > _lib.MAX_SRC:
        raise ValueError(f"the {engine} engine supports at most {_lib.MAX_SRC} sources")
    if len(sim.probe_points) > _lib.MAX_PROBE:
        raise ValueError(f"the {engine} engine supports at most {_lib.MAX_PROBE} probes")


def z_line_phase(nz, nzp, npml):
    """The line phase (floats) for field arrays with CPML on the z faces that brings
    128-byte line boundaries closest to the CPML: the sweep's z range runs from the first
    line boundary at or after k = npml+1 to the last one at or before k = nz-1-npml, and
    the cells between those and the CPML go to the per-quad shell passes.  A phase p
    puts the boundaries at k = 32m - p (p a multiple of 4: 16-byte quads).  0 when
    nz_pitch is not a multiple of 32 (rows would not share one phase) or no phase leaves
    a z range.  At nz = 1024, npml = 12: phase 16, z range [16, 1008) instead of
    [32, 992)."""
    if nzp % 32:
        return 0
    best, waste = 0, None
    for ph in range(0, 32, 4):
        kb = 32 * ((npml + 1 + ph + 31) // 32) - ph
        ke = 32 * ((nz - 1 - npml + ph) // 32) - ph
        if ke <= kb:
            continue
        w = (kb - npml - 1) + (nz - 1 - npml - ke)
        if waste is None or w < waste:
            best, waste = ph, w
    return best


def _line_phase(t, nzp):
    """Line phase (floats) of a field array as the C side reads it."""
    return (t.data_ptr() >> 2) & 31 if nzp % 32 == 0 else 0


def _offset_copy(t, offset):
    """A copy of the (contiguous) tensor t whose first element sits ``offset`` elements
    past its allocation's alignment (512 B or more)."""
    buf = torch.empty(t.numel() + offset, dtype=t.dtype, device=t.device)[offset:].view(t.shape)
    buf.copy_(t)
    return buf


def _ping_pong_sets(fields, offset=0):
    """The current and the alternate field set of the ping-pong step (6 arrays each):
    ex, ey, ez, hx, hy, hz of the current set, then of the alternate set, carved together
    by ``staggered_arrays`` (one allocation with staggered starts for arrays of 1 GiB and
    more), every array starting ``offset`` floats past a 1 KiB boundary; the current
    values are copied in.  Smaller arrays at offset 0 keep the current set and get a
    separately allocated alternate set."""
    cur = fields.ehs
    t0 = cur[0]
    if not offset and t0.numel() * t0.element_size() < _fields.ARENA_MIN_BYTES:
        return list(cur), [torch.empty_like(t) for t in cur]
    views = _fields.staggered_arrays(12, tuple(t0.shape), t0.dtype, t0.device, zero=False, offset=offset)
    with torch.cuda.stream(fields.stream):
        for v, t in zip(views[:6], cur):
            v.copy_(t)
    torch.cuda.synchronize(t0.device)
    return views[:6], views[6:]


def _check_periodic_probes(sim, axes):
    """The fused engine records probes inside its passes, before the periodic copies: no
    probe may sit on a copied cell (E tangential to a periodic axis on its last plane, H
    tangential on its first)."""
    f = sim.fields
    tangential = {0: ((1, 2), (4, 5)), 1: ((0, 2), (3, 5)), 2: ((0, 1), (3, 4))}
    for comp, pt in sim.probe_points:
        for a in axes:
            e_c, h_c = tangential[a]
            if (comp in e_c and pt[a] == f.ns[a] - 1) or (comp in h_c and pt[a] == 0):
                raise ValueError(f"a probe on a periodic copy ({pt}); use engine='kemp'")


def _check_tfsf_sources(sim, tfsf):
    """Sources and TFSF patches share the 'incident' priority; the fused engine applies the
    sources first, so no source may add to a patch cell of the same component."""
    for src in sim.sources:
        comp = _COMP[src.str_f]
        for P in (tfsf.patches_e, tfsf.patches_h):
            for q in range(P.n):
                if P.comp[q] != comp:
                    continue
                b = P.box[q]
                if all(src.pt0[a] <= b[2 * a + 1] and b[2 * a] <= src.pt1[a] for a in range(3)):
                    raise ValueError(f"a {src.str_f} source on a TFSF face cell; use engine='kemp'")


class FusedStepper:
    kernel_name = "k_fused"

    def __init__(self, sim, rows_per_thread=None, planes_per_chunk=None, step_pairs=None, z_phase=None):
        from .drude import Drude
        from .pbc import Pbc
        from .tfsf import Tfsf
        f = sim.fields
        _check_whole_step(sim, "fused", allow=(Drude, Tfsf, Pbc), cpml_axes=True)
        # periodic axes in the order the in-place engine copies them (instance order, then
        # each instance's axes)
        self.pbc_axes = [ "xyz".index(a) for i in f.instance_list if isinstance(i, Pbc) for a in i.axes]
        if len(set(self.pbc_axes)) != len(self.pbc_axes):
            raise ValueError("the fused engine copies each periodic axis once; use engine='kemp'")
        cp_ = sim.cpml
        self.cpml_axes = 0 if cp_ is None else sum(1 << a for a, d in enumerate(cp_.apply_direction) if d)
        if any((self.cpml_axes >> a) & 1 for a in self.pbc_axes):
            raise ValueError("an axis with CPML cannot be periodic")
        _check_periodic_probes(sim, self.pbc_axes)
        drudes = [i for i in f.instance_list if isinstance(i, Drude)]
        if len(drudes) > 1:
            raise ValueError("the fused engine carries one Drude box; use engine='kemp' for more")
        self.drude = drudes[0] if drudes else None
        tfsfs = [i for i in f.instance_list if isinstance(i, Tfsf)]
        if len(tfsfs) > 1:
            raise ValueError("the fused engine carries one TFSF source; use engine='kemp' for more")
        self.tfsf = tfsfs[0] if tfsfs else None
        if (self.drude is not None or self.tfsf is not None) and step_pairs:
            raise ValueError("step pairs do not support a Drude box or a TFSF source")
        if self.tfsf is not None:
            _check_tfsf_sources(sim, self.tfsf)
        cp = sim.cpml
        self.sim = sim
        npml = cp.npml if cp else 0
        # step pairs: two steps per interior sweep (needs CPML; a third field set holds the
        # first step's shell values)
        full_cpml = cp is not None and self.cpml_axes == 7 and not self.pbc_axes
        if step_pairs and not full_cpml:
            raise ValueError("step pairs need CPML on all six faces and no periodic axis")
        self.pairs = bool(full_cpml and self.drude is None and self.tfsf is None and
                          (step_pairs if step_pairs is not None else fused_pairs(f.nx, f.ny, f.nz_pitch, npml)))
        # field arrays placed so that the sweep's z range reaches close to the z CPML
        # (z_phase overrides the choice; it changes where the work goes, not the results).
        # A Drude box keeps phase 0: its z edges then stay on the sweep's 256-cell segment
        # boundaries (1024^3 with a [256, 767]^3 box, same box A/B: 16.0 ms/step at phase 0,
        # 16.4-17.0 at phase 16)
        if z_phase is None:
            z_phase = (z_line_phase(f.nz, f.nz_pitch, npml)
                       if self.cpml_axes & 4 and self.drude is None else 0)
        if z_phase % 4 or not 0 <= z_phase < 32:
            raise ValueError(f"z_phase must be a multiple of 4 in [0, 32) : {z_phase}")
        self.z_phase = z_phase
        # the arrays start (-kb) mod 256 floats past a 1 KiB boundary: line phase z_phase,
        # and k = kb (the sweep's first z cell) on a 1 KiB boundary, where its 1 KiB
        # segments start
        kb = 32 * ((npml + 1 + z_phase + 31) // 32) - z_phase
        offset = (-kb) % 256 if z_phase else 0
        f.ehs, self.alt = _ping_pong_sets(f, offset)
        if z_phase and f.ce_mat is not None and f.nz_pitch % 128 == 0:
            # CE indices (1 B per cell) placed so that the sweep's segments start on lines
            with torch.cuda.stream(f.stream):
                f.ce_mat = [_offset_copy(m, (-kb) % 128) for m in f.ce_mat]
        T, CX = fused_tiles(f.nx, f.ny, f.nz_pitch, npml, self.pairs)
        self.T = rows_per_thread or T
        self.CX = planes_per_chunk or CX
        table = f.ce_mat is not None
        if not lds_fits(self.T, self.CX, table):
            raise ValueError(f"rows_per_thread {self.T} / planes_per_chunk {self.CX} need more than 160 KiB of LDS")
        if self.pairs and not (self.T <= 9 and lds_fits_pairs(self.T, self.CX, table)):
            raise ValueError(f"rows_per_thread {self.T} / planes_per_chunk {self.CX} do not fit the two-step sweep "
                             "(at most 9 rows, 160 KiB of LDS)")
        if self.drude is not None:  # the Drude sweep runs rows+1 <= 8 waves (256 VGPRs each)
            if rows_per_thread is None:
                self.T = min(self.T, 7)
            if self.T > 7:
                raise ValueError(f"rows_per_thread {self.T}: at most 7 with a Drude box")
        # a TFSF sweep takes 13 rows per block where the grid is tall enough: its z-face
        # patches meet every row of two 256-cell segments, and at 1024^3 (CX = 64) the
        # sweep ran 11.13-11.21 ms against 11.82-11.88 ms for 9 rows (alternating in one
        # process)
        self._tfsf_rows = bool(self.tfsf is not None and rows_per_thread is None and self.drude is None
                               and self.T == 9 and f.ny >= 8 * 13
                               and lds_fits(13, self.CX, table, extra=2 * self.tfsf.n_line + 1024))
        if self._tfsf_rows:
            self.T = 13
        self.mid = (_fields.staggered_arrays(6, tuple(f.ehs[0].shape), f.ehs[0].dtype, f.device, offset=offset)
                    if self.pairs else None)
        # the Drude current of step n+1 (the two sets swap with the field sets)
        self.j_alt = [torch.zeros_like(t) for t in self.drude.J] if self.drude is not None else None
        if cp is not None:
            cp.alt_psiE()
        # CE of the sweep as one uint16 triple index per cell (rebuilt in _desc
        # whenever the material indices change)
        self._tri_kb = kb if z_phase else 0
        self.ce_tri = self.ce_tri_lut = self._tri_src = None
        self._build_ce_triples()

    def _build_ce_triples(self):
        """One uint16 index per cell into the distinct (cex, cey, cez) triples of the
        material indices (``Fields.compress_ce``), read by the sweep instead of the
        three uint8 indices (2 B per cell instead of 3; the same float values).  Skipped
        with more than MAX_CE_TRI triples or a table that does not fit the sweep's LDS (a
        TFSF sweep raised to 13 rows goes back to 9 where that lets the table fit).  Placed
        (-2 kb) mod 128 bytes past a line so that the sweep's segments start on lines."""
        f = self.sim.fields
        self.ce_tri = self.ce_tri_lut = None
        self._tri_src = f.ce_mat
        if f.ce_mat is None:
            return
        flat = [x.view(-1) for x in f.ce_mat]
        total = flat[0].numel()
        chunk = 1 << 26

        def keys(s):  # the index triple of each cell as one 24-bit key
            return (flat[0][s:s + chunk].to(torch.int32) | (flat[1][s:s + chunk].to(torch.int32) << 8) |
                    (flat[2][s:s + chunk].to(torch.int32) << 16)).long()

        with torch.cuda.stream(f.stream):
            seen = torch.zeros(1 << 24, dtype=torch.bool, device=f.device)
            for s in range(0, total, chunk):
                seen[keys(s)] = True
            vals = torch.nonzero(seen).view(-1)  # the distinct keys, ascending
            del seen
            n = vals.numel()
            # a TFSF source stages its line (2 floats per point) and two patch tables in LDS
            extra = 2 * self.tfsf.n_line + 1024 if self.tfsf is not None else 0
            if n > _lib.MAX_CE_TRI:
                return
            if not lds_fits(self.T, self.CX, True, n, extra):
                if not (self._tfsf_rows and lds_fits(9, self.CX, True, n, extra)):
                    return
                self.T = 9  # the triples fit beside the plane wave's line with the smaller tile
            lut = torch.stack([f.ce_lut[c][(vals >> (8 * c)) & 255] for c in range(3)], dim=1).contiguous()
            code = torch.zeros(1 << 24, dtype=torch.int16, device=f.device)
            code[vals] = torch.arange(n, dtype=torch.int16, device=f.device)
            off = ((-2 * self._tri_kb) % 128) // 2
            buf = torch.empty(total + off, dtype=torch.int16, device=f.device)
            tri = buf[off:]
            for s in range(0, total, chunk):
                tri[s:s + chunk] = code[keys(s)]
            del code
            self.ce_tri = tri.view(f.ce_mat[0].shape)
            self.ce_tri_lut = lut
        f.stream.synchronize()

    @property
    def bytes_per_cell(self):
        """Algorithmic HBM bytes per cell of one step (without CPML): E, H read and
        written (48 B) plus CE read as float arrays (12 B), uint8 indices (3 B) or one
        uint16 triple index (2 B)."""
        f = self.sim.fields
        if getattr(self, "ce_tri", None) is not None:
            return 50.0
        return 51.0 if f.ce_mat is not None else 60.0 if f.ce_on else 48.0

    def _desc(self):
        sim = self.sim
        f = sim.fields
        d = _lib.FusedDesc()
        d.nx, d.ny, d.nz_pitch, d.pad = f.nx, f.ny, f.nz_pitch, f.pad
        for c in range(6):
            d.e[c] = f.ehs[c].data_ptr()
            d.e_alt[c] = self.alt[c].data_ptr() if self.alt else None
        if f.ce_on:
            for c in range(3):
                d.ce[c] = f.ces[c].data_ptr()
        if f.ce_on and f.ce_mat is not None:  # CE as material indices + table
            for c in range(3):
                d.ce_mat[c] = f.ce_mat[c].data_ptr()
            d.ce_lut = f.ce_lut.data_ptr()
            if hasattr(self, "ce_tri"):
                if f.ce_mat is not self._tri_src:  # new material indices: new triples
                    self._build_ce_triples()
                if self.ce_tri is not None:
                    d.ce_tri = self.ce_tri.data_ptr()
                    d.ce_tri_lut = self.ce_tri_lut.data_ptr()
                    d.ce_tri_n = self.ce_tri_lut.shape[0]
        cp = sim.cpml
        if cp is not None:
            d.npml = cp.npml
            d.bE, d.aE, d.bH, d.aH = (cp.bE.data_ptr(), cp.aE.data_ptr(), cp.bH.data_ptr(), cp.aH.data_ptr())
            d.psiE = cp.psiE_buf.data_ptr()
            d.psiE_alt = cp.alt_psiE().data_ptr() if self.alt else None
            d.psiH = cp.psiH_buf.data_ptr()
        d.nsrc = len(sim.sources)
        for n, s in enumerate(sim.sources):
            d.src_comp[n] = _COMP[s.str_f]
            box = (s.pt0[0], s.pt1[0], s.pt0[1], s.pt1[1], s.pt0[2], s.pt1[2])
            for m in range(6):
                d.src_box[n][m] = box[m]
            d.src_sv[n] = None if s.sv is None else s.sv.data_ptr()
            d.src_svs[n] = s.svs
            d.src_overwrite[n] = int(s.is_overwrite)
        d.nprobe = len(sim.probe_points)
        for n, (c, (i, j, k)) in enumerate(sim.probe_points):
            d.probe_comp[n] = c
            d.probe_ijk[n][0], d.probe_ijk[n][1], d.probe_ijk[n][2] = i, j, k
        d.rows_per_thread = self.T
        d.planes_per_chunk = self.CX
        if self.mid:
            for c in range(6):
                d.e_mid[c] = self.mid[c].data_ptr()
        d.cpml_axes = getattr(self, "cpml_axes", 0)
        pa = getattr(self, "pbc_axes", [])
        d.npbc = len(pa)
        for q, a in enumerate(pa):
            d.pbc_axes[q] = a
        tf = getattr(self, "tfsf", None)
        if tf is not None:  # tfsf_src is set per call (its values depend on the step count)
            d.tfsf_e = tf.patches_e
            d.tfsf_h = tf.patches_h
            d.tfsf_line_e, d.tfsf_line_h, d.tfsf_abc = (tf.line_e.data_ptr(), tf.line_h.data_ptr(),
                                                        tf.abc.data_ptr())
            d.tfsf_line_n = tf.n_line
        dr = getattr(self, "drude", None)
        if dr is not None:
            box = (dr.pt0[0], dr.pt1[0], dr.pt0[1], dr.pt1[1], dr.pt0[2], dr.pt1[2])
            for m in range(6):
                d.drude_box[m] = box[m]
            for c in range(3):
                d.drude_a[c] = dr.cja[c].data_ptr()
                d.drude_b[c] = dr.cjb[c].data_ptr()
                d.drude_j[c] = dr.J[c].data_ptr()
                d.drude_j_alt[c] = self.j_alt[c].data_ptr()
        return d

    def _tvals(self, nsteps):
        srcs = self.sim.sources
        tv = np.zeros((max(nsteps, 1), max(len(srcs), 1)), np.float32)
        for n, s in enumerate(srcs):
            for k in range(nsteps):
                tv[k, n] = np.float32(s.tfunc(s.tstep + k))
        return np.ascontiguousarray(tv)

    def _call(self, nsteps, buf, timed=False):
        sim = self.sim
        f = sim.fields
        if buf is None and sim.probe_points:
            buf = torch.zeros((nsteps, len(sim.probe_points)), dtype=torch.float32, device=f.device)
        d = self._desc()
        tv = self._tvals(nsteps)
        tf = getattr(self, "tfsf", None)
        if tf is not None:  # the line's hard source per step, as Tfsf.update_e computes it
            tsrc = np.array([np.float32(np.sin(tf.wdt * (tf.tstep + k))) for k in range(max(nsteps, 1))], np.float32)
            d.tfsf_src = tsrc.ctypes.data
        parity = ctypes.c_int(0)
        args = (ctypes.byref(d), nsteps, tv.ctypes.data, None if buf is None else buf.data_ptr(),
                ctypes.byref(parity), _lib.stream_handle(f.stream))
        sweep_ms = ctypes.c_float(0.0)
        if timed:
            _lib.call("mifdtd_fused_run_timed", *args, ctypes.byref(sweep_ms))
        else:
            _lib.call("mifdtd_fused_run", *args)
        for s in sim.sources:
            s.tstep += nsteps
        if tf is not None:
            tf.tstep += nsteps
        if parity.value:
            f.ehs, self.alt = self.alt, f.ehs
            if sim.cpml is not None:
                sim.cpml.swap_psiE()
            if getattr(self, "drude", None) is not None:
                self.drude.J, self.j_alt = self.j_alt, self.drude.J
        return sweep_ms.value

    def sweep_box(self, pairs=None):
        """(xb, xe, yb, ye, kb, ke): the interior box the single-pass sweep updates (the
        cells with no CPML term; the whole grid without CPML), as the C side sets it; with
        step pairs, the two-step sweep's box (two cells further from the CPML slabs)."""
        f = self.sim.fields
        nx, ny, nz = f.nx, f.ny, f.nz
        cp = self.sim.cpml
        p = cp.npml if cp is not None else 0
        ph = _line_phase(f.ehs[0], f.nz_pitch)

        def line_up(lo):
            return 32 * ((lo + ph + 31) // 32) - ph

        def line_down(hi):
            return 32 * ((hi + ph) // 32) - ph if hi + ph >= 0 else -1

        if cp is not None and (self.pairs if pairs is None else pairs):
            box = (p + 2, nx - 2 - p, p + 2, ny - 2 - p, line_up(p + 2), line_down(nz - 2 - p))
            if box[1] - box[0] >= 4 and box[3] - box[2] >= 4 and box[5] > box[4]:
                return box
        axes = getattr(self, "cpml_axes", 7) if cp is not None else 0
        per = set(getattr(self, "pbc_axes", []))

        def bounds(a, n):  # inside the CPML slabs; a periodic axis less its last plane
            if (axes >> a) & 1:
                return p + 1, n - 1 - p
            return (0, n - 1) if a in per else (0, n)

        (xb, xe), (yb, ye) = bounds(0, nx), bounds(1, ny)
        if axes & 4:
            kb, ke = line_up(p + 1), line_down(nz - 1 - p)
        else:
            kb, ke = (0, 32 * ((nz - 1) // 32)) if 2 in per else (0, f.nz_pitch)
        if xb >= xe or yb >= ye or kb >= ke:
            return nx, nx, ny, ny, f.nz_pitch, f.nz_pitch
        return xb, xe, yb, ye, kb, ke

    def sweep_cells(self):
        xb, xe, yb, ye, kb, ke = self.sweep_box()
        return (xe - xb) * (ye - yb) * (ke - kb)

    def run(self, nsteps, buf, time_kernel=False):
        """Advance nsteps; ``buf`` is an optional device (nsteps, nprobe) probe buffer.
        With ``time_kernel`` HIP events bracket every interior-sweep launch on the fields'
        stream and the average sweep duration (ms) is returned."""
        if not time_kernel:
            self._call(nsteps, buf)
            return None
        return self._call(nsteps, buf, timed=True)


class SlabStepper(FusedStepper):
    """In-place slab-ordered whole step: per slab of ``slab_planes`` x planes, an E pass
    then an H pass, each with CPML, sources and probes fused in."""
    kernel_name = "k_slab"

    def __init__(self, sim, slab_planes=None):
        f = sim.fields
        _check_whole_step(sim, "slab")
        self.sim = sim
        self.alt = None
        self.mid = None
        self.pairs = False
        self.drude = None
        self.tfsf = None
        self.pbc_axes = []
        self.cpml_axes = 0 if sim.cpml is None else 7
        self.T = self.CX = 0
        cp = sim.cpml
        self.S = slab_planes or slab_planes_for(f, cp)

    @property
    def bytes_per_cell(self):
        """Algorithmic HBM bytes per cell of one step (without CPML), as the fused
        engine's: CE as uint8 indices 3 B, as float arrays 12 B."""
        f = self.sim.fields
        return 51.0 if f.ce_on and f.ce_mat is not None else 60.0 if f.ce_on else 48.0

    def run(self, nsteps, buf, time_kernel=False):
        """Advance nsteps; ``buf`` is an optional device (nsteps, nprobe) probe buffer.
        With ``time_kernel`` each step's launches are bracketed by HIP events on the
        fields' stream and the average step time (ms) is returned."""
        if not time_kernel:
            self._call(nsteps, buf)
            return None
        st = self.sim.fields.stream
        evs = []
        for k in range(nsteps):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(st)
            self._call(1, None if buf is None else buf[k:k + 1])
            b.record(st)
            evs.append((a, b))
        torch.cuda.synchronize()
        return sum(a.elapsed_time(b) for a, b in evs) / max(nsteps, 1)

    def _call(self, nsteps, buf, timed=False):
        sim = self.sim
        f = sim.fields
        if buf is None and sim.probe_points:
            buf = torch.zeros((nsteps, len(sim.probe_points)), dtype=torch.float32, device=f.device)
        d = self._desc()
        tv = self._tvals(nsteps)
        _lib.call("mifdtd_slab_run", ctypes.byref(d), self.S, nsteps, tv.ctypes.data,
                  None if buf is None else buf.data_ptr(), _lib.stream_handle(f.stream))
        for s in sim.sources:
            s.tstep += nsteps


def slab_planes_for(f, cp):
    return slab_planes(f.nx, f.ny, f.nz_pitch, cp.npml if cp else 0, f.ce_on)
