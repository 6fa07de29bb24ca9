"""Builds ``libmifdtd.so`` (all HIP kernels + the C ABI) in-tree for gfx950, and the
KEMP core program libraries over it.

``python -m fdtd_accelerate_amd.build`` compiles every ``csrc/*.hip`` to an object in
``_build/`` (in parallel) and links the shared library next to this file.  Flags:
  -ffp-contract=off   keep the reference's mul/add rounding (no FMA contraction)
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed-FP32 VALU forms
It then compiles ``csrc/kemp_program.c`` once per KEMP program (precision, coeff_use,
pad), as K7/cpu/core.py:52-85 generates one per combination, into
``kemp/libkemp_<precision>_<coeff>_pad<p>.so``, and once more per combination in the
split form of K7/cpu/core_split.py (``..._pad<p>_split.so``), each linked to libmifdtd.so.
Last it compiles ``csrc/h5io.cpp`` (host code: HDF5 snapshot and geometry files) with the
host C++ compiler against the HDF5 C library under ``$MIFDTD_HDF5`` (default
``/opt/conda``) into ``libmifdtd_h5.so``.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libmifdtd.so")
KEMP_DIR = os.path.join(PKG, "kemp")
CC = os.environ.get("CC", shutil.which("gcc") or "gcc")
CXX = os.environ.get("CXX", shutil.which("g++") or "g++")
H5_LIB = os.path.join(PKG, "libmifdtd_h5.so")
HDF5_ROOT = os.environ.get("MIFDTD_HDF5", "/opt/conda")

# (precision_float, pads) of the reference's programs: single pads nz to 4 floats, double
# to 2 (K7/cpu/fields.py:27-32 with the 16-byte SSE width).
KEMP_PRECISIONS = {"single": (0, 1, 2, 3), "double": (0, 1)}
KEMP_COEFFS = ("", "e", "h", "eh")


def kemp_program_path(precision_float: str, coeff_use: str, pad: int, split: bool = False) -> str:
    """Path of the program library of (precision_float, coeff_use, pad), split form or not."""
    return os.path.join(KEMP_DIR, f"libkemp_{precision_float}_{coeff_use or 'none'}_pad{pad}"
                                  f"{'_split' if split else ''}.so")
INCLUDE = os.path.join(os.path.dirname(PKG), "include")

ARCH = os.environ.get("MIFDTD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(os.path.basename(src))[0] + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    """Compile all sources and link ``libmifdtd.so``; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    if not srcs:
        raise RuntimeError(f"no HIP sources under {CSRC}")
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    progs = build_kemp_programs()
    h5 = build_h5()
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources and {len(progs)} KEMP programs"
              + (f"; {h5}" if h5 else f"; no HDF5 under {HDF5_ROOT}: {H5_LIB} not built"), file=sys.stderr)
    return LIB


def build_h5():
    """Compile ``libmifdtd_h5.so`` (HDF5 snapshot and geometry files); returns its path,
    or None when the HDF5 C library is not installed (the snapshot module then raises)."""
    inc, lib = os.path.join(HDF5_ROOT, "include"), os.path.join(HDF5_ROOT, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        return None
    cmd = [CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra", "-pthread", f"-I{INCLUDE}", f"-I{inc}",
           os.path.join(CSRC, "h5io.cpp"), "-o", H5_LIB + ".tmp", f"-L{lib}", "-lhdf5", "-lz", f"-Wl,-rpath,{lib}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 or r.stderr.strip():
        raise RuntimeError(f"h5io.cpp failed:\n{r.stderr}")
    os.replace(H5_LIB + ".tmp", H5_LIB)
    return H5_LIB


def _compile_program(args):
    prec, coeff, pad, split = args
    out = kemp_program_path(prec, coeff, pad, split)
    cmd = [CC, "-O2", "-std=c99", "-fPIC", "-shared", "-fvisibility=hidden", f"-I{INCLUDE}",
           f"-DKEMP_DOUBLE={int(prec == 'double')}", f"-DKEMP_CE={int('e' in coeff)}",
           f"-DKEMP_CH={int('h' in coeff)}", f"-DKEMP_PAD={pad}", f"-DKEMP_SPLIT={int(split)}",
           os.path.join(CSRC, "kemp_program.c"), "-o", out + ".tmp",
           f"-L{PKG}", "-l:libmifdtd.so", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"KEMP program {prec}/{coeff or 'none'}/pad{pad}/split{int(split)} failed:\n{r.stderr}")
    os.replace(out + ".tmp", out)
    return out


def build_kemp_programs() -> list:
    """Compile every KEMP program library (needs libmifdtd.so built)."""
    os.makedirs(KEMP_DIR, exist_ok=True)
    jobs = [(p, c, pad, sp) for p, pads in KEMP_PRECISIONS.items() for c in KEMP_COEFFS for pad in pads
            for sp in (False, True)]
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        return list(ex.map(_compile_program, jobs))


if __name__ == "__main__":
    build(verbose=True)
