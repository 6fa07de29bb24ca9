"""ORACLE — test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker (or, for bench.py, as the
timed CPU baseline).  The product package ``fdtd_accelerate_amd`` never imports it.

What it holds:

* ``lib`` — ctypes handle on ``oracle/liboracle.so`` (``make -C oracle``), the scalar C
  restatement of the reference hot path in ``oracle_body.h`` (float and double).
* ``naive_update_e/h`` — a NumPy slice-form restatement of ``K7/naive/core.py:24-47``
  (a second, independent formulation of the core update used to pin the C one).
* ``ref_core(dtype, pad)`` — the reference's own KEMP C core compiled in place by
  ``make -C oracle ref`` into ``oracle/_ref`` (only where /root/reference exists).
* ``cpml_tables`` / ``pulse`` — closed forms of the CPML profile and the
  Gaussian-modulated source (``KU/cpu/cpml.py:107-126``, ``KB/source/source.py:247-249``).
* ``OracleSim`` — the full step (E core, Drude currents, CPML-E faces, E sources, TFSF
  E faces, periodic E planes, H core, CPML-H faces, H sources, TFSF H faces, periodic H
  planes, probes) in the order of ``KU/cpu/test/020-dielectric-cpml.py:58-73`` and the
  instance priorities of ``K7/util/common.py:383-397``.
* ``OracleS4Sim`` and ``s4_*`` — the 4th-order symplectic stages of
  ``S4/Python-C/dielectric.c:5-96`` in the stage loop of ``S4/PyCUDA/100-base.py:124-141``
  (the C stages are pinned against the Python restatement and the integrator's
  closed-form amplification matrix), with the CPML this engine adds to that path.
* ``pbc_copy`` — the periodic-boundary plane copy of ``K7/cpu/pbc.py:25-31`` with the
  plane table of ``K7/util/common_exchange.py:1-29``.  Parity pinned by the replication
  invariant (a P-periodic state stays P-periodic: a run over period P equals, bit for
  bit, the same state replicated over period 2P), which a wrong plane index, copy
  direction or component fails.
* ``OracleTfsf`` — the normal-incidence plane wave of ``KU/dim3/tfsf.py:14-242``.
  Pinned by the scattered-field invariant (in vacuum the field outside the box stays at
  round-off level), which a flipped coefficient or a correction one plane off fails.
* ``drude_coefficients`` / ``OracleDrude`` — the auxiliary-current update of
  ``KU/dim3/cpu/core/drude_ade.c:14-99`` with the coefficients of
  ``KU/dim3/cpu/matter.py:192-202``.  Pinned by the discrete plasma oscillation of a
  uniform periodic medium and its damping factor.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(HERE, "liboracle.so")
_REF_DIR = os.path.join(HERE, "_ref")

_P = ctypes.c_void_p
_I = ctypes.c_int


def build(ref: bool = False) -> None:
    """Build liboracle.so (and, with ``ref``, the reference core in oracle/_ref)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _load():
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    for sfx, real in (("f32", ctypes.c_float), ("f64", ctypes.c_double)):
        getattr(lib, f"ora_update_e_{sfx}").argtypes = [_I] * 4 + [_P] * 9
        getattr(lib, f"ora_update_h_{sfx}").argtypes = [_I] * 4 + [_P] * 9
        getattr(lib, f"ora_cpml_e_{sfx}").argtypes = [_I] * 7 + [_P] * 13
        getattr(lib, f"ora_cpml_h_{sfx}").argtypes = [_I] * 7 + [_P] * 13
        getattr(lib, f"ora_src_{sfx}").argtypes = [_P] + [_I] * 9 + [real, _P, real, _I]
        getattr(lib, f"ora_s4_update_h_{sfx}").argtypes = [_I] * 3 + [real] + [_P] * 9
        getattr(lib, f"ora_s4_update_e_{sfx}").argtypes = [_I] * 3 + [real] + [_P] * 9
        getattr(lib, f"ora_s4_cpml_e_{sfx}").argtypes = [_I] * 6 + [real] + [_P] * 13
        getattr(lib, f"ora_s4_cpml_h_{sfx}").argtypes = [_I] * 6 + [real] + [_P] * 13
        for n in ("update_e", "update_h", "cpml_e", "cpml_h", "src", "s4_update_h", "s4_update_e",
                  "s4_cpml_e", "s4_cpml_h"):
            getattr(lib, f"ora_{n}_{sfx}").restype = None
    return lib


lib = _load()


def _sfx(a: np.ndarray) -> str:
    if a.dtype == np.float32:
        return "f32"
    if a.dtype == np.float64:
        return "f64"
    raise TypeError(f"oracle arrays must be float32 or float64, got {a.dtype}")


def _ptr(a):
    if a is None:
        return None
    assert a.flags.c_contiguous or a.flags.f_contiguous
    return a.ctypes.data


# ----------------------------------------------------------------------------- core
def update_e(ehs, ces, pad: int) -> None:
    """Core E update in place (``K7/cpu/src/core.c:11-55``); ``ces`` may be None (0.5)."""
    ex, ey, ez, hx, hy, hz = ehs
    nx, ny, nzp = ex.shape
    cs = ces if ces is not None else (None, None, None)
    getattr(lib, f"ora_update_e_{_sfx(ex)}")(nx, ny, nzp, pad, *[_ptr(a) for a in ehs],
                                              *[_ptr(c) for c in cs])


def update_h(ehs, chs, pad: int) -> None:
    """Core H update in place (``K7/cpu/src/core.c:59-105``); ``chs`` may be None (0.5)."""
    ex = ehs[0]
    nx, ny, nzp = ex.shape
    cs = chs if chs is not None else (None, None, None)
    getattr(lib, f"ora_update_h_{_sfx(ex)}")(nx, ny, nzp, pad, *[_ptr(a) for a in ehs],
                                              *[_ptr(c) for c in cs])


def naive_update_e(ehs, ces, pad: int) -> None:
    """NumPy slice form of ``K7/naive/core.py:24-34`` plus the zeroed lanes of the C core
    (``K7/cpu/src/core.c:31,41,51``: ex/ey lanes k >= nz-1 and ez pad lanes become 0)."""
    ex, ey, ez, hx, hy, hz = ehs
    nzp = ex.shape[2]
    nz = nzp - pad
    dt = ex.dtype
    cex, cey, cez = ces if ces is not None else [np.full(ex.shape, 0.5, dt)] * 3
    z1 = slice(0, nz - 1)
    z0 = slice(1, nz)
    zz = slice(0, nz)
    ex[:, :-1, z1] += cex[:, :-1, z1] * ((hz[:, 1:, z1] - hz[:, :-1, z1]) - (hy[:, :-1, z0] - hy[:, :-1, z1]))
    ey[:-1, :, z1] += cey[:-1, :, z1] * ((hx[:-1, :, z0] - hx[:-1, :, z1]) - (hz[1:, :, z1] - hz[:-1, :, z1]))
    ez[:-1, :-1, zz] += cez[:-1, :-1, zz] * ((hy[1:, :-1, zz] - hy[:-1, :-1, zz]) - (hx[:-1, 1:, zz] - hx[:-1, :-1, zz]))
    ex[:, :-1, nz - 1:] = 0
    ey[:-1, :, nz - 1:] = 0
    ez[:-1, :-1, nz:] = 0


def naive_update_h(ehs, chs, pad: int) -> None:
    """NumPy slice form of ``K7/naive/core.py:37-47`` plus the zeroed lanes of the C core
    (``K7/cpu/src/core.c:79-80,90-91,101``: hx/hy lane k=0 and pad lanes become 0)."""
    ex, ey, ez, hx, hy, hz = ehs
    nzp = ex.shape[2]
    nz = nzp - pad
    dt = ex.dtype
    chx, chy, chz = chs if chs is not None else [np.full(ex.shape, 0.5, dt)] * 3
    z0 = slice(1, nz)
    z1 = slice(0, nz - 1)
    zz = slice(0, nz)
    hx[:, 1:, z0] -= chx[:, 1:, z0] * ((ez[:, 1:, z0] - ez[:, :-1, z0]) - (ey[:, 1:, z0] - ey[:, 1:, z1]))
    hy[1:, :, z0] -= chy[1:, :, z0] * ((ex[1:, :, z0] - ex[1:, :, z1]) - (ez[1:, :, z0] - ez[:-1, :, z0]))
    hz[1:, 1:, zz] -= chz[1:, 1:, zz] * ((ey[1:, 1:, zz] - ey[:-1, 1:, zz]) - (ex[1:, 1:, zz] - ex[1:, :-1, zz]))
    hx[:, 1:, 0] = 0
    hy[1:, :, 0] = 0
    hx[:, 1:, nz:] = 0
    hy[1:, :, nz:] = 0
    hz[1:, 1:, nz:] = 0


# --------------------------------------------------------------- reference KEMP core
def ref_available() -> bool:
    return os.path.isdir("/root/reference") or os.path.isdir(_REF_DIR)


def ref_core(dtype, pad: int, split: bool = False):
    """The reference's ``update_e/update_h`` (``K7/cpu/src/core.c``; with ``split`` the
    ``K7/cpu/src/core_split.c`` form taking idx0, nmax after nz) built in place with the
    masks of the given pad; ``coeff_use='eh'`` signatures (6 fields + 3 coefficient arrays)."""
    sfx = "f32" if np.dtype(dtype) == np.float32 else "f64"
    path = os.path.join(_REF_DIR, f"libkemp_{'split' if split else 'core'}_{sfx}_pad{pad}.so")
    if not os.path.exists(path):
        build(ref=True)
    cl = ctypes.CDLL(path)
    for n in ("update_e", "update_h"):
        fn = getattr(cl, n)
        fn.argtypes = [_I] * (5 if split else 3) + [_P] * 9
        fn.restype = None
    return cl


# ------------------------------------------------------------------------- CPML
def cpml_sigma_max(npml: int, m: int = 4, R: float = 1e-6, eta: float = 1.0, dx: float = 1.0) -> float:
    """Textbook polynomial-grading sigma_max for a target normal reflection R."""
    return -(m + 1) * math.log(R) / (2.0 * eta * npml * dx)


def cpml_tables(npml: int, m: int = 4, R: float = 1e-6, dt: float = 0.5, dtype=np.float32):
    """Closed-form CPML tables in the 2(npml+1) layout of ``KU/cpu/cpml.py:107-126``.

    sigma_E front[p+1] = smax*((npml-0.5-p)/npml)^m, back[npml+1+p] = smax*((0.5+p)/npml)^m;
    sigma_H front[p+1] = smax*((npml-p)/npml)^m,    back[npml+2+p] = smax*((1+p)/npml)^m.
    kappa = 1, alpha = 0 (normalized units, eps0 = 1):  b = exp(-sigma*dt), a = b - 1
    where sigma > 0 and a = 0 where sigma = 0.  Computed in float64, cast to ``dtype``.
    """
    smax = cpml_sigma_max(npml, m, R)
    sE = np.zeros(2 * (npml + 1))
    sH = np.zeros(2 * (npml + 1))
    for p in range(npml):
        sE[p + 1] = smax * ((npml - 0.5 - p) / npml) ** m
        sE[npml + 1 + p] = smax * ((0.5 + p) / npml) ** m
        sH[p + 1] = smax * ((npml - p) / npml) ** m
        sH[npml + 2 + p] = smax * ((1.0 + p) / npml) ** m
    bE = np.exp(-sE * dt)
    bH = np.exp(-sH * dt)
    aE = np.where(sE > 0, bE - 1.0, 0.0)
    aH = np.where(sH > 0, bH - 1.0, 0.0)
    return tuple(np.ascontiguousarray(t, dtype=dtype) for t in (bE, aE, bH, aH))


def psi_shape(axis: int, nx: int, ny: int, nzp: int, npml: int):
    return [(npml, ny, nzp), (nx, npml, nzp), (nx, ny, npml)][axis]


def cpml_e(axis, side, ehs, ces, psi1, psi2, b, a, pad, npml):
    ex = ehs[0]
    nx, ny, nzp = ex.shape
    cs = ces if ces is not None else (None, None, None)
    getattr(lib, f"ora_cpml_e_{_sfx(ex)}")(axis, side, nx, ny, nzp, pad, npml,
                                            *[_ptr(x) for x in ehs], *[_ptr(c) for c in cs],
                                            _ptr(psi1), _ptr(psi2), _ptr(b), _ptr(a))


def cpml_h(axis, side, ehs, chs, psi1, psi2, b, a, pad, npml):
    ex = ehs[0]
    nx, ny, nzp = ex.shape
    cs = chs if chs is not None else (None, None, None)
    getattr(lib, f"ora_cpml_h_{_sfx(ex)}")(axis, side, nx, ny, nzp, pad, npml,
                                            *[_ptr(x) for x in ehs], *[_ptr(c) for c in cs],
                                            _ptr(psi1), _ptr(psi2), _ptr(b), _ptr(a))


# ----------------------------------------------------------------------- sources
def pulse(t: float, t0: float = 100.0, width: float = 30.0, period: float = 20.0) -> float:
    """Gaussian-modulated sinusoid of ``KB/source/source.py:247-249`` in step units:
    exp(-0.5*((t-t0)/width)^2) * cos(2*pi*(t-t0)/period), evaluated in float64."""
    return math.exp(-0.5 * ((t - t0) / width) ** 2) * math.cos(2.0 * math.pi * (t - t0) / period)


def src_add(f, pt0, pt1, tval, svalue, overwrite=False):
    nx, ny, nzp = f.shape
    real = f.dtype.type
    sv = svalue if isinstance(svalue, np.ndarray) else None
    svs = 0.0 if sv is not None else float(real(svalue))
    if sv is not None:
        sv = np.ascontiguousarray(sv, dtype=f.dtype)
    getattr(lib, f"ora_src_{_sfx(f)}")(_ptr(f), nx, ny, nzp, pt0[0], pt1[0], pt0[1], pt1[1],
                                        pt0[2], pt1[2], float(real(tval)), _ptr(sv), svs,
                                        int(overwrite))


# --------------------------------------------------------------------- 4th order
def s4_update_h(ehs, chs, cl):
    ex = ehs[0]
    nx, ny, nz = ex.shape
    assert ex.flags.f_contiguous
    getattr(lib, f"ora_s4_update_h_{_sfx(ex)}")(nx, ny, nz, float(ex.dtype.type(cl)),
                                                 *[_ptr(a) for a in ehs], *[_ptr(c) for c in chs])


def s4_update_e(ehs, ces, dl):
    ex = ehs[0]
    nx, ny, nz = ex.shape
    assert ex.flags.f_contiguous
    getattr(lib, f"ora_s4_update_e_{_sfx(ex)}")(nx, ny, nz, float(ex.dtype.type(dl)),
                                                 *[_ptr(a) for a in ehs], *[_ptr(c) for c in ces])


# 5-stage symplectic integrator of S4/PyCUDA/100-base.py:60-63 (H weights c, E weights
# d; the step closes with H(c[4]) = H(c1), 100-base.py:138-141).
S4_C = (0.17399689146541, -0.12038504121430, 0.89277629949778, -0.12038504121430, 0.17399689146541)
S4_D = (0.62337932451322, -0.12337932451322, -0.12337932451322, 0.62337932451322, 0.0)


def s4_psi_shape(axis, nx, ny, nz, npml):
    """psi layout of the 4th-order CPML (x fastest): F-order shapes."""
    return [(npml, ny, nz), (nx, npml, nz), (nx, ny, npml)][axis]


def s4_cpml_e(axis, side, ehs, ces, psi1, psi2, b, a, w, npml):
    ex = ehs[0]
    nx, ny, nz = ex.shape
    getattr(lib, f"ora_s4_cpml_e_{_sfx(ex)}")(axis, side, nx, ny, nz, npml, float(ex.dtype.type(w)),
                                               *[_ptr(x) for x in ehs], *[_ptr(c) for c in ces],
                                               _ptr(psi1), _ptr(psi2), _ptr(b), _ptr(a))


def s4_cpml_h(axis, side, ehs, chs, psi1, psi2, b, a, w, npml):
    ex = ehs[0]
    nx, ny, nz = ex.shape
    getattr(lib, f"ora_s4_cpml_h_{_sfx(ex)}")(axis, side, nx, ny, nz, npml, float(ex.dtype.type(w)),
                                               *[_ptr(x) for x in ehs], *[_ptr(c) for c in chs],
                                               _ptr(psi1), _ptr(psi2), _ptr(b), _ptr(a))


class OracleS4Sim:
    """CPU oracle of the 4th-order symplectic step (``S4/PyCUDA/100-base.py:124-141``):
      for s in 0..3: H(c_s) -> CPML-H(c_s) -> E(d_s) -> CPML-E(d_s) -> E sources at t+s/5
      then H(c_4) -> CPML-H(c_4); probes after the step.
    Fields are (nx, ny, nz) Fortran-order arrays.  ``ces``/``chs`` hold S/(24 eps_r) /
    S/24 (``100-base.py:75``); default vacuum.  CPML tables of stage s are those of
    the time step w_s*S (``cpml_tables(dt=w_s*S)``, w_s in float64).  Sources follow the
    reference's stage timing: after E stage s the source adds
    float32(tfunc(tstep + s/5)) * spatial_value (``100-base.py:136``), tstep from 1.
    """

    def __init__(self, nx, ny, nz, S, npml=0, dtype=np.float32, ces=None, chs=None, cpml_R=1e-6, cpml_m=4):
        self.nx, self.ny, self.nz, self.S, self.npml = nx, ny, nz, S, npml
        self.dtype = dtype
        self.ehs = [np.zeros((nx, ny, nz), dtype, order="F") for _ in range(6)]
        c0 = dtype(S / 24.0)
        self.ces = [np.asfortranarray(c, dtype=dtype) for c in ces] if ces is not None else \
            [np.full((nx, ny, nz), c0, dtype, order="F") for _ in range(3)]
        self.chs = [np.asfortranarray(c, dtype=dtype) for c in chs] if chs is not None else \
            [np.full((nx, ny, nz), c0, dtype, order="F") for _ in range(3)]
        self.cl = np.array(S4_C, dtype)
        self.dl = np.array(S4_D, dtype)
        self.sources = []
        self.probes = []
        self.series = []
        self.tstep = 1
        if npml:
            self.tabE = [cpml_tables(npml, cpml_m, cpml_R, dt=S4_D[s] * S, dtype=dtype)[:2] for s in range(4)]
            self.tabH = [cpml_tables(npml, cpml_m, cpml_R, dt=S4_C[s] * S, dtype=dtype)[2:] for s in range(5)]
            self.psiE, self.psiH = {}, {}
            for ax in range(3):
                for sd in range(2):
                    shp = s4_psi_shape(ax, nx, ny, nz, npml)
                    self.psiE[ax, sd] = tuple(np.zeros(shp, dtype, order="F") for _ in range(2))
                    self.psiH[ax, sd] = tuple(np.zeros(shp, dtype, order="F") for _ in range(2))

    def field(self, name):
        return self.ehs[("ex", "ey", "ez", "hx", "hy", "hz").index(name)]

    def add_source(self, str_f, pt0, pt1, tfunc, svalue=1.0):
        if str_f[0] != "e":
            raise ValueError("4th-order sources drive E components")
        self.sources.append((str_f, tuple(pt0), tuple(pt1), tfunc, svalue))

    def add_probe(self, str_f, pt):
        self.probes.append((str_f, tuple(pt)))

    def _h(self, s):
        s4_update_h(self.ehs, self.chs, self.cl[s])
        if self.npml:
            b, a = self.tabH[s]
            for ax in range(3):
                for sd in range(2):
                    p1, p2 = self.psiH[ax, sd]
                    s4_cpml_h(ax, sd, self.ehs, self.chs, p1, p2, b, a, self.cl[s], self.npml)

    def _e(self, s):
        s4_update_e(self.ehs, self.ces, self.dl[s])
        if self.npml:
            b, a = self.tabE[s]
            for ax in range(3):
                for sd in range(2):
                    p1, p2 = self.psiE[ax, sd]
                    s4_cpml_e(ax, sd, self.ehs, self.ces, p1, p2, b, a, self.dl[s], self.npml)
        real = self.dtype
        for str_f, p0, p1, tfunc, sv in self.sources:
            f = self.field(str_f)
            box = (slice(p0[0], p1[0] + 1), slice(p0[1], p1[1] + 1), slice(p0[2], p1[2] + 1))
            tv = real(tfunc(self.tstep + s / 5.0))
            f[box] += tv * (np.asarray(sv, real) if isinstance(sv, np.ndarray) else real(sv))

    def step(self):
        for s in range(4):
            self._h(s)
            self._e(s)
        self._h(4)
        self.tstep += 1
        if self.probes:
            self.series.append([self.field(n)[p] for n, p in self.probes])

    def run(self, nsteps):
        for _ in range(nsteps):
            self.step()


# --------------------------------------------------------------------- full step
# Periodic boundaries (K7/util/common_exchange.py:1-29): the field components tangential
# to each axis; after the E half step plane n-1 takes plane 0, after the H half step
# plane 0 takes plane n-1 (n the real cell count; the x and y planes include the z pad).
PBC_COMPONENTS = {"x": {"e": (1, 2), "h": (4, 5)}, "y": {"e": (0, 2), "h": (3, 5)},
                  "z": {"e": (0, 1), "h": (3, 4)}}


def pbc_copy(ehs, ns, axes, e_or_h):
    """K7/cpu/pbc.py:25-31: for each axis of ``axes`` in order, copy the tangential
    components' boundary plane across the grid (ns = real (nx, ny, nz))."""
    for axis in axes:
        a = "xyz".index(axis)
        n = ns[a]
        src, dst = (0, n - 1) if e_or_h == "e" else (n - 1, 0)
        for c in PBC_COMPONENTS[axis][e_or_h]:
            f = ehs[c]
            get = [slice(None)] * 3
            put = [slice(None)] * 3
            get[a], put[a] = src, dst
            f[tuple(put)] = f[tuple(get)]


class OracleTfsf:
    """Total-field / scattered-field plane wave, normal incidence (``KU/dim3/tfsf.py``,
    kuFDTD v0.2.0; its oblique branch is empty).  A float32 1-D grid (``Fdtd1d``,
    ``:14-40``) carries the wave: E[1:-1] -= 0.5 (H[2:] - H[1:-1]), two-step-delay
    boundaries at both ends, a hard source E[3] = sin(wdt * tstep); H[1:] -= 0.5 (E[1:] -
    E[:-1]).  1-D index 5 is the box's first plane along the propagation axis.  After
    each 3-D half step the faces of the box [pt0, pt1] (inclusive) named by
    ``apply_direction`` take the incident correction of ``calc_face_list`` /
    ``calc_inc_list`` (``:138-226``): E planes just outside a face += 0.5 * c * H_inc, H
    planes on the face += 0.5 * c * E_inc, c = +-sin/cos of the polarization angle.
    Parity is pinned by the scattered-field invariant: in vacuum the field outside the
    box stays at round-off level while the box carries the wave.
    """

    def __init__(self, pt0, pt1, apply_direction, wdt, axis, p_angle):
        self.pt0 = list(pt0)
        self.pt1 = [v + 1 for v in pt1]   # tfsf.py:111 (pt2 += 1)
        self.apply = apply_direction
        self.wdt = wdt
        self.axis = axis
        a = "xyz".index(axis)
        n = self.pt1[a] - self.pt0[a] + 1
        self.E = np.zeros(n + 10, np.float32)
        self.H = np.zeros(n + 10, np.float32)
        self.abc = [np.float32(0)] * 4
        self.tstep = 1
        i1, i2 = 5, self.pt1[a] - self.pt0[a] + 5
        # float32 coefficients: the reference multiplies float32 arrays by them under
        # value-based casting, i.e. in float32
        si, co = np.float32(np.sin(p_angle)), np.float32(np.cos(p_angle))
        sl1, sl2, sl3, sl4, sl5 = (slice(i1 - 1, i1), slice(i1, i1 + 1), slice(i1, i2), slice(i1, i2 + 1),
                                   slice(i2, i2 + 1))
        # (ce1, ce2, ch1, ch2, sliceE1, sliceE2, sliceH1, sliceH2) of Inc1..Inc6
        z = np.float32(0)
        inc = [(-si, co, co, si, sl1, sl1, sl2, sl2), (si, -co, -co, -si, sl5, sl5, sl5, sl5),
               (z, si, z, -co, sl4, sl3, sl3, sl4), (z, -si, z, co, sl4, sl3, sl3, sl4),
               (-co, z, -si, z, sl3, sl4, sl4, sl3), (co, z, si, z, sl3, sl4, sl4, sl3)]
        order, newaxis = {"x": ([0, 1, 2, 3, 4, 5], [0, 0, 1, 1, 1, 1]),
                          "y": ([4, 5, 0, 1, 2, 3], [1, 1, 0, 0, 0, 0]),
                          "z": ([2, 3, 4, 5, 0, 1], [0, 0, 0, 0, 0, 0])}[axis]
        (i1_, j1, k1), (i2_, j2, k2) = self.pt0, self.pt1
        slx, slxp, sly, slyp, slz, slzp = (slice(i1_, i2_), slice(i1_, i2_ + 1), slice(j1, j2), slice(j1, j2 + 1),
                                           slice(k1, k2), slice(k1, k2 + 1))
        # (e1, e2, h1, h2 component indices, slice_e1, slice_e2, slice_h1, slice_h2) per face
        faces = [((1, 2, 4, 5), (i1_ - 1, slyp, slz), (i1_ - 1, sly, slzp), (i1_, sly, slzp), (i1_, slyp, slz)),
                 ((1, 2, 4, 5), (i2_, slyp, slz), (i2_, sly, slzp), (i2_, sly, slzp), (i2_, slyp, slz)),
                 ((2, 0, 5, 3), (slx, j1 - 1, slzp), (slxp, j1 - 1, slz), (slxp, j1, slz), (slx, j1, slzp)),
                 ((2, 0, 5, 3), (slx, j2, slzp), (slxp, j2, slz), (slxp, j2, slz), (slx, j2, slzp)),
                 ((0, 1, 3, 4), (slxp, sly, k1 - 1), (slx, slyp, k1 - 1), (slx, slyp, k1), (slxp, sly, k1)),
                 ((0, 1, 3, 4), (slxp, sly, k2), (slx, slyp, k2), (slx, slyp, k2), (slxp, sly, k2))]
        self.pairs = []
        for f in range(6):
            if ("f" if f % 2 == 0 else "b") in apply_direction[f // 2]:
                self.pairs.append((faces[f], inc[order[f]], newaxis[f]))

    @staticmethod
    def _expand(v, newaxis_i):
        sl = [slice(None)]
        sl.insert(newaxis_i, np.newaxis)
        return v[tuple(sl)]

    def update_e(self, ehs):
        E, H = self.E, self.H
        E[1:-1] -= 0.5 * (H[2:] - H[1:-1])
        E[0] = self.abc[0]
        self.abc[0] = self.abc[1]
        self.abc[1] = E[1]
        E[-1] = self.abc[2]
        self.abc[2] = self.abc[3]
        self.abc[3] = E[-2]
        E[3] = np.float32(np.sin(self.wdt * self.tstep))
        self.tstep += 1
        for (comps, se1, se2, sh1, sh2), (ce1, ce2, ch1, ch2, sE1, sE2, sH1, sH2), nax in self.pairs:
            h1 = self._expand(ch1 * H[sH1], nax)
            h2 = self._expand(ch2 * H[sH2], nax)
            ehs[comps[0]][se1] += 0.5 * h1
            ehs[comps[1]][se2] += 0.5 * h2

    def update_h(self, ehs):
        E, H = self.E, self.H
        H[1:] -= 0.5 * (E[1:] - E[:-1])
        for (comps, se1, se2, sh1, sh2), (ce1, ce2, ch1, ch2, sE1, sE2, sH1, sH2), nax in self.pairs:
            e1 = self._expand(ce1 * E[sE1], nax)
            e2 = self._expand(ce2 * E[sE2], nax)
            ehs[comps[2]][sh1] += 0.5 * e1
            ehs[comps[3]][sh2] += 0.5 * e2


def drude_coefficients(plasma_freq, gamma, dt, dtype=np.float32):
    """CJA = (2 - gamma dt) / (2 + gamma dt), CJB = 4 (dt wp)^2 / (2 + gamma dt)
    (``KU/dim3/cpu/matter.py:192-202``), in the arrays' precision."""
    g = np.asarray(gamma, dtype)
    w = np.asarray(plasma_freq, dtype)
    tmp = 2 + g * dt
    return (2 - g * dt) / tmp, 4 * (dt * w) ** 2 / tmp


class OracleDrude:
    """Drude dispersion by the auxiliary-current method (``KU/dim3/cpu/core/drude_ade.c``,
    ``KU/dim3/cpu/matter.py:104-221``): right after the core E update, on every cell of
    the box [pt0, pt1] (inclusive) and each component c,
        e' = e - CE_c J_c ;  J_c' = CJA_c J_c + CJB_c e'
    (J, CJA, CJB region-sized; CE the grid's E coefficient, 0.5 without arrays).
    Priority 'current' (after core E, before CPML-E).  Pinned by the plasma oscillation
    of a uniform periodic medium: with zero curl the update is the 2x2 recurrence of
    determinant CJA and trace 1 + CJA - CJB CE, so a uniform E oscillates at
    arccos(1 - (wp dt)^2 / 2) per step (gamma = 0) and decays by sqrt(CJA) per step.
    """

    def __init__(self, pt0, pt1, cja, cjb, dtype=np.float32):
        self.sl = tuple(slice(a, b + 1) for a, b in zip(pt0, pt1))
        shp = tuple(b - a + 1 for a, b in zip(pt0, pt1))
        self.cja = [np.broadcast_to(np.asarray(c, dtype), shp).copy() for c in cja]
        self.cjb = [np.broadcast_to(np.asarray(c, dtype), shp).copy() for c in cjb]
        self.J = [np.zeros(shp, dtype) for _ in range(3)]

    def update_e(self, ehs, ces):
        for c in range(3):
            e = ehs[c][self.sl]
            ce = ces[c][self.sl] if ces is not None else e.dtype.type(0.5)
            e_tmp = e - ce * self.J[c]
            ehs[c][self.sl] = e_tmp
            self.J[c] = self.cja[c] * self.J[c] + self.cjb[c] * e_tmp


class OracleSim:
    """CPU oracle of one Yee+CPML simulation on the KEMP layout.

    Step order (``KU/cpu/test/020-dielectric-cpml.py:58-73``; instance priorities
    core < pml < incident, ``K7/util/common.py:383-397``):
      core E -> CPML-E (x f, x b, y f, y b, z f, z b) -> E sources -> core H ->
      CPML-H (same face order) -> H sources -> probes (read after H, ``K7/main.py:303``).
    Sources follow ``K7/cpu/incident_direct.py:75-77``: tstep starts at 1 and is
    incremented after each use; the value is REAL(tfunc(tstep)) * spatial_value.
    """

    def __init__(self, nx, ny, nz, pad=0, npml=0, dtype=np.float32, ces=None, chs=None,
                 cpml_faces=((True, True), (True, True), (True, True)), cpml_R=1e-6, cpml_m=4, pbc=""):
        self.nx, self.ny, self.nz, self.pad = nx, ny, nz, pad
        self.pbc = pbc  # periodic axes, applied in this order
        self.tfsf = []  # OracleTfsf instances (priority 'incident', after the sources)
        self.drude = []  # OracleDrude instances (priority 'current', after the core E update)
        self.nzp = nz + pad
        self.dtype = np.dtype(dtype)
        shp = (nx, ny, self.nzp)
        self.ehs = [np.zeros(shp, dtype) for _ in range(6)]
        self.ces = None if ces is None else [np.ascontiguousarray(c, dtype) for c in ces]
        self.chs = None if chs is None else [np.ascontiguousarray(c, dtype) for c in chs]
        self.npml = npml
        self.faces = cpml_faces
        self.sources = []   # (str_f, pt0, pt1, tfunc, svalue, overwrite, [tstep])
        self.probes = []    # (str_f, (i,j,k))
        self.series = []
        if npml > 0:
            self.bE, self.aE, self.bH, self.aH = cpml_tables(npml, cpml_m, cpml_R, dtype=dtype)
            self.psiE = {}
            self.psiH = {}
            for ax in range(3):
                for sd in range(2):
                    if self.faces[ax][sd]:
                        s = psi_shape(ax, nx, ny, self.nzp, npml)
                        self.psiE[ax, sd] = (np.zeros(s, dtype), np.zeros(s, dtype))
                        self.psiH[ax, sd] = (np.zeros(s, dtype), np.zeros(s, dtype))

    _names = ("ex", "ey", "ez", "hx", "hy", "hz")

    def field(self, name):
        return self.ehs[self._names.index(name)]

    def add_source(self, str_f, pt0, pt1, tfunc, svalue=1.0, overwrite=False):
        self.sources.append([str_f, tuple(pt0), tuple(pt1), tfunc, svalue, overwrite, 1])

    def add_probe(self, str_f, pt):
        self.probes.append((str_f, tuple(pt)))

    def _sources(self, e_or_h):
        for s in self.sources:
            if s[0][0] != e_or_h:
                continue
            src_add(self.field(s[0]), s[1], s[2], s[3](s[6]), s[4], s[5])
            s[6] += 1

    def step(self):
        update_e(self.ehs, self.ces, self.pad)
        for d in self.drude:
            d.update_e(self.ehs, self.ces)
        if self.npml:
            for ax in range(3):
                for sd in range(2):
                    if (ax, sd) in self.psiE:
                        p1, p2 = self.psiE[ax, sd]
                        cpml_e(ax, sd, self.ehs, self.ces, p1, p2, self.bE, self.aE, self.pad, self.npml)
        self._sources("e")
        for t in self.tfsf:
            t.update_e(self.ehs)
        pbc_copy(self.ehs, (self.nx, self.ny, self.nz), self.pbc, "e")
        update_h(self.ehs, self.chs, self.pad)
        if self.npml:
            for ax in range(3):
                for sd in range(2):
                    if (ax, sd) in self.psiH:
                        p1, p2 = self.psiH[ax, sd]
                        cpml_h(ax, sd, self.ehs, self.chs, p1, p2, self.bH, self.aH, self.pad, self.npml)
        self._sources("h")
        for t in self.tfsf:
            t.update_h(self.ehs)
        pbc_copy(self.ehs, (self.nx, self.ny, self.nz), self.pbc, "h")
        if self.probes:
            self.series.append([self.field(n)[p] for n, p in self.probes])

    def run(self, nsteps):
        for _ in range(nsteps):
            self.step()
