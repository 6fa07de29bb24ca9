"""ctypes binding of ``libmifdtd.so`` (the C ABI declared in ``include/mifdtd.h``).

There is no CPU fallback: importing a compute entry point without the built library
raises ``RuntimeError`` (build it with ``python -m fdtd_accelerate_amd.build``).
"""
from __future__ import annotations

import ctypes
import os
import threading

from .build import LIB as LIB_PATH

_I = ctypes.c_int
_F = ctypes.c_float
_P = ctypes.c_void_p
_LL = ctypes.c_longlong


class CpmlHalf(ctypes.Structure):
    """``mifdtd_cpml_half``: npml, device b/a tables, psi1[6], psi2[6] (index 2*axis+side)."""
    _fields_ = [("npml", _I), ("b", _P), ("a", _P), ("psi1", _P * 6), ("psi2", _P * 6)]


MAX_SRC = 4
MAX_PROBE = 8
TFSF_MAX_PATCH = 12
MAX_CE_TRI = 4096  # MIFDTD_FUSED_MAX_CE_TRI


class TfsfPatches(ctypes.Structure):
    """``mifdtd_tfsf_patches``: face patches of one TFSF half step."""
    _fields_ = [("n", _I), ("axis", _I), ("comp", _I * TFSF_MAX_PATCH), ("box", (_I * 6) * TFSF_MAX_PATCH),
                ("base", _I * TFSF_MAX_PATCH), ("coef", ctypes.c_float * TFSF_MAX_PATCH),
                ("start", ctypes.c_uint * (TFSF_MAX_PATCH + 1))]


class FusedDesc(ctypes.Structure):
    """``mifdtd_fused_desc`` (include/mifdtd.h)."""
    _fields_ = [
        ("nx", _I), ("ny", _I), ("nz_pitch", _I), ("pad", _I),
        ("e", _P * 6), ("e_alt", _P * 6), ("ce", _P * 3),
        ("npml", _I), ("bE", _P), ("aE", _P), ("bH", _P), ("aH", _P),
        ("psiE", _P), ("psiE_alt", _P), ("psiH", _P),
        ("nsrc", _I), ("src_comp", _I * MAX_SRC), ("src_box", (_I * 6) * MAX_SRC),
        ("src_sv", _P * MAX_SRC), ("src_svs", _F * MAX_SRC), ("src_overwrite", _I * MAX_SRC),
        ("nprobe", _I), ("probe_comp", _I * MAX_PROBE), ("probe_ijk", (_I * 3) * MAX_PROBE),
        ("rows_per_thread", _I), ("planes_per_chunk", _I),
        ("ce_mat", _P * 3), ("ce_lut", _P), ("e_mid", _P * 6),
        ("drude_box", _I * 6), ("drude_a", _P * 3), ("drude_b", _P * 3), ("drude_j", _P * 3),
        ("drude_j_alt", _P * 3),
        ("tfsf_e", TfsfPatches), ("tfsf_h", TfsfPatches), ("tfsf_line_e", _P), ("tfsf_line_h", _P),
        ("tfsf_abc", _P), ("tfsf_line_n", _I), ("tfsf_src", _P),
        ("cpml_axes", _I), ("npbc", _I), ("pbc_axes", _I * 3),
        ("ce_tri", _P), ("ce_tri_lut", _P), ("ce_tri_n", _I),
    ]


class InstanceDesc(ctypes.Structure):
    """``mifdtd_instance`` (include/mifdtd.h)."""
    _fields_ = [
        ("nx", _I), ("ny", _I), ("nz_pitch", _I), ("pad", _I),
        ("e", _P * 6), ("ce", _P * 3), ("npml", _I), ("psiE", _P), ("psiH", _P),
        ("nsrc", _I), ("src_comp", _I * MAX_SRC), ("src_box", (_I * 6) * MAX_SRC),
        ("src_sv", _P * MAX_SRC), ("src_svs", _F * MAX_SRC), ("src_overwrite", _I * MAX_SRC),
        ("nprobe", _I), ("probe_comp", _I * MAX_PROBE), ("probe_ijk", (_I * 3) * MAX_PROBE),
        ("ce_mat", _P * 3), ("ce_lut", _P),
    ]


class DesignBox(ctypes.Structure):
    """``mifdtd_design_box``."""
    _fields_ = [("i0", _I), ("j0", _I), ("k0", _I), ("di", _I), ("dj", _I), ("dk", _I)]


# name -> (restype, argtypes)
_SIGS = {
    "mifdtd_last_error": (ctypes.c_char_p, []),
    "mifdtd_version": (_I, []),
    "mifdtd_update_e": (_I, [_I] * 4 + [_P] * 9 + [_P]),
    "mifdtd_update_h": (_I, [_I] * 4 + [_P] * 9 + [_P]),
    "mifdtd_update_e_f64": (_I, [_I] * 4 + [_P] * 9 + [_P]),
    "mifdtd_update_h_f64": (_I, [_I] * 4 + [_P] * 9 + [_P]),
    "mifdtd_update_e_range": (_I, [_I] * 4 + [_LL, _LL] + [_P] * 9 + [_P]),
    "mifdtd_update_h_range": (_I, [_I] * 4 + [_LL, _LL] + [_P] * 9 + [_P]),
    "mifdtd_update_e_f64_range": (_I, [_I] * 4 + [_LL, _LL] + [_P] * 9 + [_P]),
    "mifdtd_update_h_f64_range": (_I, [_I] * 4 + [_LL, _LL] + [_P] * 9 + [_P]),
    "mifdtd_update_e_lut": (_I, [_I] * 4 + [_P] * 10 + [_P]),
    "mifdtd_cpml_face_x_e": (_I, [_I] * 6 + [_P] * 10 + [_P]),
    "mifdtd_cpml_face_x_h": (_I, [_I] * 6 + [_P] * 8 + [_P]),
    "mifdtd_cpml_face_y_e": (_I, [_I] * 6 + [_P] * 10 + [_P]),
    "mifdtd_cpml_face_y_h": (_I, [_I] * 6 + [_P] * 8 + [_P]),
    "mifdtd_cpml_face_z_e": (_I, [_I] * 6 + [_P] * 10 + [_P]),
    "mifdtd_cpml_face_z_h": (_I, [_I] * 6 + [_P] * 8 + [_P]),
    "mifdtd_cpml_e": (_I, [_I] * 4 + [ctypes.POINTER(CpmlHalf)] + [_P] * 9 + [_P]),
    "mifdtd_cpml_h": (_I, [_I] * 4 + [ctypes.POINTER(CpmlHalf)] + [_P] * 9 + [_P]),
    "mifdtd_source": (_I, [_P] + [_I] * 9 + [_F, _P, _F, _I, _P]),
    "mifdtd_gather": (_I, [_P, _P, _I, _P, _P]),
    "mifdtd_pbc": (_I, [_P, _P] + [_I] * 6 + [_P]),
    "mifdtd_drude_e": (_I, [_P] * 5 + [_I] * 9 + [_P]),
    "mifdtd_tfsf_line": (_I, [_P, _P, _P, _I, _I, _F, _P]),
    "mifdtd_tfsf_faces": (_I, [_P, _P, _P, _I, _I, _I, _I, _P]),
    "mifdtd_fused_run": (_I, [ctypes.POINTER(FusedDesc), _I, _P, _P, ctypes.POINTER(_I), _P]),
    "mifdtd_fused_run_timed": (_I, [ctypes.POINTER(FusedDesc), _I, _P, _P, ctypes.POINTER(_I), _P,
                                    ctypes.POINTER(ctypes.c_float)]),
    "mifdtd_slab_run": (_I, [ctypes.POINTER(FusedDesc), _I, _I, _P, _P, _P]),
    "mifdtd_batch_create": (_I, [ctypes.POINTER(InstanceDesc), _I, _P, _P, _P, _P, ctypes.POINTER(_P)]),
    "mifdtd_batch_sizes": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_uint)]),
    "mifdtd_batch_run": (_I, [_P, _I, _P, _P, _I, _I, _P]),
    "mifdtd_batch_destroy": (_I, [_P]),
    "mifdtd_batch_run_static": (_I, [ctypes.POINTER(InstanceDesc), _I, _P, _P, _P, _P, _I, _P, _P, _I, _P]),
    "mifdtd_adjoint_step": (_I, [_I] * 4 + [ctypes.POINTER(_P * 6), ctypes.POINTER(_P * 3), _I, _P, _P,
                                            _P, _P, _P, _P, _P, _I, ctypes.POINTER(DesignBox), _P, _P, _P]),
    "mifdtd_design_curl": (_I, [_I] * 4 + [_P] * 3 + [ctypes.POINTER(DesignBox), _P, _P]),
    "s4_update_h": (None, [_I] * 3 + [_F] + [_P] * 9),
    "s4_update_e": (None, [_I] * 3 + [_F] + [_P] * 9),
    "mifdtd_s4_update_h": (_I, [_I] * 3 + [_F] + [_P] * 9 + [_F, _P]),
    "mifdtd_s4_update_e": (_I, [_I] * 3 + [_F] + [_P] * 9 + [_F, _P]),
    "mifdtd_s4_update_e_lut": (_I, [_I] * 3 + [_F] + [_P] * 10 + [_P]),
    "mifdtd_s4_cpml_e": (_I, [_I] * 3 + [ctypes.POINTER(CpmlHalf), _F] + [_P] * 9 + [_F, _P]),
    "mifdtd_s4_cpml_h": (_I, [_I] * 3 + [ctypes.POINTER(CpmlHalf), _F] + [_P] * 9 + [_F, _P]),
}

# Symbols declared in include/mifdtd.h that the library must export.
EXPORTS = tuple(_SIGS)

_lock = threading.Lock()
_lib = None


def load():
    """Load (once) and return the ctypes library; raise if it is not built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is not built: the HIP path has no fallback. "
                    "Run `python -m fdtd_accelerate_amd.build` (or __graft_entry__.build()).")
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in _SIGS.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def last_error() -> str:
    msg = load().mifdtd_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed (rc={rc}): {last_error()}")


def call(name: str, *args) -> None:
    """Call a status-returning entry point and raise RuntimeError on failure."""
    check(getattr(load(), name)(*args), name)


def stream_handle(stream=None):
    """Raw hipStream_t of a torch stream (default: torch's current stream)."""
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)
