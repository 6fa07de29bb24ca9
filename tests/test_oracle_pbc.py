"""Periodic boundaries in the oracle (``K7/cpu/pbc.py:25-31``, plane table
``K7/util/common_exchange.py:1-29``), pinned by the replication invariant.

With periodic axis a the last plane (E) / first plane (H) of the tangential components is
a copy of the opposite one, so the state has period P = n_a - 1.  A P-periodic state
stays P-periodic, so a run on a grid of n_a = P + 1 and a run on a grid of n_a = 2P + 1
started from the same state replicated twice must agree bit for bit, every component,
every cell (each cell of the larger grid evaluates the same expression on the same
values).  A copy in the wrong direction, to the wrong plane or of the wrong components
breaks that; the last test plants each of these in the oracle and checks that the
invariant fails."""
import numpy as np
import pytest

from oracle import oracle as O


def _replicate(r, axes, reps, pad):
    """Tile the periodic block ``r`` (real cells, shape P per periodic axis) to
    reps * P + 1 cells along each periodic axis (the +1 plane repeats plane 0), pad z."""
    out = r
    for a in range(3):
        if "xyz"[a] in axes:
            n = r.shape[a]
            out = np.take(out, np.arange(reps * n + 1) % n, axis=a)
    if pad:
        out = np.pad(out, ((0, 0), (0, 0), (0, pad)))
    return np.ascontiguousarray(out, np.float32)


def _run(axes, block, reps, steps, seed, npml, pad, ce):
    rng = np.random.default_rng(seed)
    fields = [rng.standard_normal(block).astype(np.float32) for _ in range(6)]
    ces = [(0.5 / rng.uniform(1, 4, block)).astype(np.float32) for _ in range(3)] if ce else None
    faces = tuple((False, False) if "xyz"[a] in axes else (True, True) for a in range(3))
    out = []
    for r in (1, reps):
        ehs = [_replicate(f, axes, r, pad) for f in fields]
        nx, ny, nzp = ehs[0].shape
        cs = None if ces is None else [_replicate(c, axes, r, pad) for c in ces]
        o = O.OracleSim(nx, ny, nzp - pad, pad, npml, np.float32, ces=cs, cpml_faces=faces, pbc=axes)
        for k in range(6):
            o.ehs[k][...] = ehs[k]
        o.run(steps)
        out.append(o)
    return out


CASES = [("x", (6, 5, 7), 0, 0, True), ("y", (5, 7, 6), 0, 4, True), ("z", (6, 5, 9), 0, 4, False),
         ("xy", (5, 6, 7), 0, 0, True), ("zx", (6, 5, 8), 0, 4, True), ("xyz", (5, 4, 6), 0, 2, True),
         ("x", (8, 14, 18), 3, 2, True)]


@pytest.mark.parametrize("axes,block,npml,pad,ce", CASES)
def test_oracle_pbc_replication_invariant(axes, block, npml, pad, ce):
    small, big = _run(axes, block, 2, 6, 7, npml, pad, ce)
    for k in range(6):
        a = small.ehs[k]
        b = big.ehs[k]
        # every period of the larger run equals the smaller run
        for shift in range(2):
            sa = [slice(None)] * 3
            sb = [slice(None)] * 3
            for ax in range(3):
                if "xyz"[ax] in axes:
                    sa[ax] = slice(0, block[ax] + 1)
                    sb[ax] = slice(shift * block[ax], shift * block[ax] + block[ax] + 1)
            assert np.ascontiguousarray(b[tuple(sb)]).tobytes() == np.ascontiguousarray(a[tuple(sa)]).tobytes(), \
                (axes, k, shift)
    # the state moved and the boundary planes are live
    assert not np.array_equal(small.ehs[2], _replicate(np.zeros(block, np.float32), axes, 1, pad))


def test_oracle_pbc_copy_table():
    """The plane table: E copies plane 0 to plane n-1, H plane n-1 to plane 0, of the
    two tangential components only, z planes at the real nz - 1 (not the pad)."""
    ns = (4, 5, 6)
    shp = (4, 5, 8)
    for axis, comps in (("x", (1, 2)), ("y", (0, 2)), ("z", (0, 1))):
        a = "xyz".index(axis)
        for half, base in (("e", 0), ("h", 3)):
            ehs = [np.arange(np.prod(shp), dtype=np.float32).reshape(shp) + 1000 * c for c in range(6)]
            before = [x.copy() for x in ehs]
            O.pbc_copy(ehs, ns, axis, half)
            def idx(p):
                return tuple(p if d == a else slice(None) for d in range(3))
            for c in range(6):
                if c in (base + comps[0], base + comps[1]):
                    src, dst = (0, ns[a] - 1) if half == "e" else (ns[a] - 1, 0)
                    assert np.array_equal(ehs[c][idx(dst)], before[c][idx(src)])
                    rest = np.ones(shp, bool)
                    rest[idx(dst)] = False
                    assert np.array_equal(ehs[c][rest], before[c][rest])
                else:
                    assert np.array_equal(ehs[c], before[c]), (axis, half, c)


def test_oracle_pbc_invariant_catches_planted_mistakes(monkeypatch):
    good = O.pbc_copy

    def wrong_direction(ehs, ns, axes, e_or_h):
        good(ehs, ns, axes, "h" if e_or_h == "e" else "e")

    def off_by_one(ehs, ns, axes, e_or_h):
        good(ehs, tuple(n - 1 for n in ns), axes, e_or_h)

    def wrong_components(ehs, ns, axes, e_or_h):
        for axis in axes:
            a = "xyz".index(axis)
            n = ns[a]
            src, dst = (0, n - 1) if e_or_h == "e" else (n - 1, 0)
            for c in ((0, 1) if e_or_h == "e" else (3, 4)):  # the x-axis mistake: normal + one tangential
                sl_g = [slice(None)] * 3
                sl_p = [slice(None)] * 3
                sl_g[a], sl_p[a] = src, dst
                ehs[c][tuple(sl_p)] = ehs[c][tuple(sl_g)]

    for bad in (wrong_direction, off_by_one, wrong_components):
        monkeypatch.setattr(O, "pbc_copy", bad)
        small, big = _run("x", (6, 5, 7), 2, 5, 3, 0, 0, True)
        same = all(big.ehs[k][:7].tobytes() == small.ehs[k].tobytes() and big.ehs[k][6:].tobytes() == small.ehs[k].tobytes()
                   for k in range(6))
        assert not same, bad.__name__
    monkeypatch.setattr(O, "pbc_copy", good)
