"""Reference CPU baseline (test infrastructure, CPU only; never runs on the GPU box).

Times the reference's own KEMP C/SSE/OpenMP core (K7/cpu/src/core.c, built in place by
``make -C oracle ref`` as the coeff_use='e' program K7/cpu/core.py:52-85 generates) driving
the plain Yee step, with the CPML cost added by this directory's OpenMP C restatement of
kuFDTD's CPML (the reference's own plain-C CPML, 2nd_order-Yee/CUDA/fdtd3d_011_cpu_cpml.c,
includes hdf5.h, which this image lacks, so it is not built).  One step is the reference
order: update_e, CPML-E on the six faces, the Ez point source, update_h, CPML-H.

    python oracle/time_reference.py            # both workloads, as built and with FTZ/DAZ
    python oracle/time_reference.py --n 128 --npml 10 --steps 1000 [--ftz]

Each (workload, FTZ) run is a separate process: OpenMP worker threads take the MXCSR of
the thread that creates them, so FTZ/DAZ is set before the first parallel region.
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# run as a script, this directory is sys.path[0] and would shadow the oracle package
sys.path = [os.path.dirname(HERE)] + [p for p in sys.path if os.path.abspath(p or ".") != HERE]

from oracle import oracle as O  # noqa: E402

_I, _P = ctypes.c_int, ctypes.c_void_p
E_PROGRAM = os.path.join(HERE, "_ref", "libkemp_core_f32e_pad0.so")


def load_e_program():
    if not os.path.exists(E_PROGRAM):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)
    lib = ctypes.CDLL(E_PROGRAM)
    lib.update_e.argtypes = [_I] * 3 + [_P] * 9
    lib.update_h.argtypes = [_I] * 3 + [_P] * 6
    lib.update_e.restype = lib.update_h.restype = None
    return lib


def check_e_program(lib):
    """The coeff_use='e' program equals the all-array program with CH = 0.5 (bit for bit)."""
    rng = np.random.default_rng(3)
    shp = (6, 7, 16)
    a = [rng.random(shp, dtype=np.float32) for _ in range(6)]
    ce = [rng.random(shp, dtype=np.float32) for _ in range(3)]
    b = [x.copy() for x in a]
    for _ in range(3):
        lib.update_e(*shp, *[x.ctypes.data for x in a + ce])
        lib.update_h(*shp, *[x.ctypes.data for x in a])
        O.update_e(b, ce, 0)
        O.update_h(b, None, 0)
    assert all(x.tobytes() == y.tobytes() for x, y in zip(a, b)), "coeff 'e' program differs from the oracle"


def run(n, npml, steps, ftz):
    """One timed run in this process: returns a dict of timings."""
    O.lib.ora_set_ftz(int(ftz))
    lib = load_e_program()
    check_e_program(lib)
    f = [np.zeros((n, n, n), np.float32) for _ in range(6)]
    ce = [np.full((n, n, n), 0.5, np.float32) for _ in range(3)]  # vacuum (the arithmetic is the same for any CE)
    bE, aE, bH, aH = O.cpml_tables(npml)
    psiE = {(ax, sd): [np.zeros(O.psi_shape(ax, n, n, n, npml), np.float32) for _ in range(2)]
            for ax in range(3) for sd in range(2)}
    psiH = {k: [np.zeros_like(v[0]), np.zeros_like(v[1])] for k, v in psiE.items()}
    c = n // 2
    ptr = [x.ctypes.data for x in f]
    cptr = [x.ctypes.data for x in ce]
    t_core = t_pml = 0.0
    wall0 = time.perf_counter()
    for s in range(steps):
        t0 = time.perf_counter()
        lib.update_e(n, n, n, *ptr, *cptr)
        t1 = time.perf_counter()
        for (ax, sd), (p1, p2) in psiE.items():
            O.cpml_e(ax, sd, f, ce, p1, p2, bE, aE, 0, npml)
        t2 = time.perf_counter()
        f[2][c, c, c] += np.float32(O.pulse(s + 1))
        t3 = time.perf_counter()
        lib.update_h(n, n, n, *ptr)
        t4 = time.perf_counter()
        for (ax, sd), (p1, p2) in psiH.items():
            O.cpml_h(ax, sd, f, None, p1, p2, bH, aH, 0, npml)
        t5 = time.perf_counter()
        t_core += (t1 - t0) + (t4 - t3)
        t_pml += (t2 - t1) + (t5 - t4)
    wall = time.perf_counter() - wall0
    cells = n ** 3
    return {"n": n, "npml": npml, "steps": steps, "ftz": bool(ftz), "threads": int(os.environ.get("OMP_NUM_THREADS", "0")),
            "wall_s": round(wall, 3), "core_s": round(t_core, 3), "cpml_s": round(t_pml, 3),
            "gcells_step": round(cells * steps / wall / 1e9, 4), "gcells_core": round(cells * steps / t_core / 1e9, 4),
            "probe_ez": float(f[2][min(c + n // 4, n - 1), c, c])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int)
    ap.add_argument("--npml", type=int)
    ap.add_argument("--steps", type=int)
    ap.add_argument("--ftz", action="store_true")
    ap.add_argument("--threads", type=int, default=8)
    a = ap.parse_args()
    if a.n:
        print(json.dumps(run(a.n, a.npml, a.steps, a.ftz)), flush=True)
        return
    env = dict(os.environ, OMP_NUM_THREADS=str(a.threads))
    for n, npml, steps in ((128, 10, 1000), (1024, 12, 5)):
        for ftz in (False, True):
            cmd = [sys.executable, os.path.abspath(__file__), "--n", str(n), "--npml", str(npml), "--steps", str(steps)]
            if ftz:
                cmd.append("--ftz")
            subprocess.run(cmd, env=env, check=True)


if __name__ == "__main__":
    main()
