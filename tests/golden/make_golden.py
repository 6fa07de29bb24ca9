"""Writes tests/golden/kemp_core.npz (float) and kemp_core_f64.npz (double): golden
vectors of the reference's own KEMP core.

Run in the build container (needs /root/reference): it compiles
K7/cpu/src/core.c in place (``make -C oracle ref``), drives it through ctypes on seeded
random inputs, and stores inputs and outputs.  The parameter grid follows the
reference's test (``K7/cpu/test/test_core.py:69-89``: ufunc e/h, nz 61..64, coeff_use
'', 'e', 'h', tmax 1 and 10) on smaller grids, plus alternating E/H steps; one case
at the reference's own grid (40, 50, 64, tmax 10, coeff 'e') is stored as a SHA-256
digest of the output with its input seed.

The double file follows the same grid with the reference's 'double' programs (2-lane
SSE2, nz padded to 2: nz 13 and 14 give pads 1 and 0).

    python tests/golden/make_golden.py            # float
    python tests/golden/make_golden.py --f64      # double
    python tests/golden/make_golden.py --split    # split form (core_split.c), float and double
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


def rand_fields(rng, nx, ny, nz, nzp, n, dtype=np.float32):
    out = []
    for _ in range(n):
        a = np.zeros((nx, ny, nzp), dtype)
        a[:, :, :nz] = rng.random((nx, ny, nz), dtype=dtype)
        out.append(a)
    return out


def run_ref(pad, ehs, ces, chs, ufunc, tmax):
    ref = O.ref_core(ehs[0].dtype, pad)
    nx, ny, nzp = ehs[0].shape
    a = [x.copy() for x in ehs]
    for _ in range(tmax):
        if ufunc in ("e", "eh"):
            ref.update_e(nx, ny, nzp, *[x.ctypes.data for x in a], *[c.ctypes.data for c in ces])
        if ufunc in ("h", "eh"):
            ref.update_h(nx, ny, nzp, *[x.ctypes.data for x in a], *[c.ctypes.data for c in chs])
    return a


def coeffs(rng, coeff_use, nx, ny, nz, nzp, dtype=np.float32):
    half = lambda: np.full((nx, ny, nzp), 0.5, dtype)
    ces = rand_fields(rng, nx, ny, nz, nzp, 3, dtype) if "e" in coeff_use else [half() for _ in range(3)]
    chs = rand_fields(rng, nx, ny, nz, nzp, 3, dtype) if "h" in coeff_use else [half() for _ in range(3)]
    for c in ces + chs:
        c[:, :, nz:] = 0.5
    return ces, chs


def main():
    O.build(ref=True)
    rec = {}
    nx, ny = 5, 6
    case = 0
    for nz in (13, 14, 15, 16):
        nzp = (nz + 3) // 4 * 4
        pad = nzp - nz
        for ufunc, coeff_use, tmax in (("e", "", 1), ("e", "e", 1), ("h", "", 1), ("h", "h", 1),
                                       ("eh", "eh", 5)):
            rng = np.random.default_rng(1000 + case)
            if ufunc == "e":
                ehs = [np.zeros((nx, ny, nzp), np.float32) for _ in range(3)] + rand_fields(rng, nx, ny, nz, nzp, 3)
            elif ufunc == "h":
                ehs = rand_fields(rng, nx, ny, nz, nzp, 3) + [np.zeros((nx, ny, nzp), np.float32) for _ in range(3)]
            else:
                ehs = rand_fields(rng, nx, ny, nz, nzp, 6)
            ces, chs = coeffs(rng, coeff_use, nx, ny, nz, nzp)
            out = run_ref(pad, ehs, ces, chs, ufunc, tmax)
            p = f"c{case}_"
            rec[p + "meta"] = np.array([nx, ny, nz, nzp, pad, tmax, {"e": 0, "h": 1, "eh": 2}[ufunc],
                                        int("e" in coeff_use), int("h" in coeff_use)], np.int64)
            rec[p + "in"] = np.stack(ehs)
            if "e" in coeff_use:
                rec[p + "ce"] = np.stack(ces)
            if "h" in coeff_use:
                rec[p + "ch"] = np.stack(chs)
            rec[p + "out"] = np.stack(out)
            case += 1
    rec["ncases"] = np.array(case)

    # the reference test's own grid, stored as a digest (K7/cpu/test/test_core.py:86)
    nx, ny, nz = 40, 50, 64
    seed = 4242
    rng = np.random.default_rng(seed)
    ehs = [np.zeros((nx, ny, nz), np.float32) for _ in range(3)] + rand_fields(rng, nx, ny, nz, nz, 3)
    ces = rand_fields(rng, nx, ny, nz, nz, 3)
    chs = [np.full((nx, ny, nz), 0.5, np.float32) for _ in range(3)]
    out = run_ref(0, ehs, ces, chs, "e", 10)
    h = hashlib.sha256(b"".join(o.tobytes() for o in out[:3])).hexdigest()
    rec["big_seed"] = np.array(seed)
    rec["big_sha256"] = np.array(h)
    rec["big_probe"] = np.array([out[0][20, 25, 30], out[1][10, 5, 3], out[2][39 - 1, 48, 63]], np.float32)
    np.savez_compressed(os.path.join(HERE, "kemp_core.npz"), **rec)
    print("wrote", case, "cases; big digest", h)


def main_f64():
    """Golden vectors of the reference's double programs (pads 0 and 1)."""
    O.build(ref=True)
    rec = {}
    nx, ny = 5, 6
    case = 0
    f8 = np.float64
    for nz in (13, 14, 15, 16):
        nzp = (nz + 1) // 2 * 2
        pad = nzp - nz
        for ufunc, coeff_use, tmax in (("e", "", 1), ("e", "e", 1), ("h", "", 1), ("h", "h", 1),
                                       ("eh", "eh", 5)):
            rng = np.random.default_rng(5000 + case)
            zeros = lambda: np.zeros((nx, ny, nzp), f8)
            if ufunc == "e":
                ehs = [zeros() for _ in range(3)] + rand_fields(rng, nx, ny, nz, nzp, 3, f8)
            elif ufunc == "h":
                ehs = rand_fields(rng, nx, ny, nz, nzp, 3, f8) + [zeros() for _ in range(3)]
            else:
                ehs = rand_fields(rng, nx, ny, nz, nzp, 6, f8)
            ces, chs = coeffs(rng, coeff_use, nx, ny, nz, nzp, f8)
            out = run_ref(pad, ehs, ces, chs, ufunc, tmax)
            p = f"c{case}_"
            rec[p + "meta"] = np.array([nx, ny, nz, nzp, pad, tmax, {"e": 0, "h": 1, "eh": 2}[ufunc],
                                        int("e" in coeff_use), int("h" in coeff_use)], np.int64)
            rec[p + "in"] = np.stack(ehs)
            if "e" in coeff_use:
                rec[p + "ce"] = np.stack(ces)
            if "h" in coeff_use:
                rec[p + "ch"] = np.stack(chs)
            rec[p + "out"] = np.stack(out)
            case += 1
    rec["ncases"] = np.array(case)
    np.savez_compressed(os.path.join(HERE, "kemp_core_f64.npz"), **rec)
    print("wrote", case, "double cases")


def big_inputs(seed):
    """Inputs of the digest case (shared with the tests)."""
    nx, ny, nz = 40, 50, 64
    rng = np.random.default_rng(int(seed))
    ehs = [np.zeros((nx, ny, nz), np.float32) for _ in range(3)] + rand_fields(rng, nx, ny, nz, nz, 3)
    ces = rand_fields(rng, nx, ny, nz, nz, 3)
    return ehs, ces


def split_ranges(nx, ny, nzp, vec):
    """Element ranges of the split-form cases: the whole grid, K7/cpu/core_split.py's
    'pre' / 'post' ranges of E and H, a range starting and ending mid-row with nmax off
    a vector boundary, the last vector, and an empty range."""
    nyz, n = ny * nzp, nx * ny * nzp
    assert nx >= 3 and n // 2 + 1 > vec * 7  # every range inside the grid (the reference loop has no bounds check)
    return [(0, n), (nyz, 3 * nyz), (0, nyz), (0, 2 * nyz), (2 * nyz, 3 * nyz),
            (vec * 7, n // 2 + 1), (n - vec, n), (vec * 3, vec * 3)]


def main_split():
    """Golden vectors of the reference's split programs (core_split.c): one E and one H
    call over each range of ``split_ranges``, float pads 3 and 0, double pad 1; the output
    holds the three updated components (the other three are unchanged inputs)."""
    O.build(ref=True)
    rec = {}
    nx, ny = 3, 4
    case = 0
    for dtype, vec, nzs in ((np.float32, 4, (13, 16)), (np.float64, 2, (13,))):
        for nz in nzs:
            nzp = (nz + vec - 1) // vec * vec
            pad = nzp - nz
            ref = O.ref_core(dtype, pad, split=True)
            for idx0, nmax in split_ranges(nx, ny, nzp, vec):
                for ufunc in ("e", "h"):
                    rng = np.random.default_rng(9000 + case)
                    ehs = rand_fields(rng, nx, ny, nz, nzp, 6, dtype)
                    ces, chs = coeffs(rng, "eh", nx, ny, nz, nzp, dtype)
                    a = [x.copy() for x in ehs]
                    fn, cs = (ref.update_e, ces) if ufunc == "e" else (ref.update_h, chs)
                    fn(nx, ny, nzp, idx0, nmax, *[x.ctypes.data for x in a], *[c.ctypes.data for c in cs])
                    p = f"c{case}_"
                    rec[p + "meta"] = np.array([nx, ny, nz, nzp, pad, idx0, nmax, int(ufunc == "h"),
                                                int(dtype == np.float64)], np.int64)
                    rec[p + "in"] = np.stack(ehs)
                    rec[p + "coef"] = np.stack(cs)
                    rec[p + "out"] = np.stack(a[:3] if ufunc == "e" else a[3:])
                    case += 1
    rec["ncases"] = np.array(case)
    np.savez_compressed(os.path.join(HERE, "kemp_core_split.npz"), **rec)
    print("wrote", case, "split cases")


if __name__ == "__main__":
    if sys.argv[1:] == ["--f64"]:
        main_f64()
    elif sys.argv[1:] == ["--split"]:
        main_split()
    else:
        main()
