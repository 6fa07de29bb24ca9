"""The oracle's TFSF plane wave (``kuFDTD/v0.2.0/dim3/tfsf.py``, normal incidence),
pinned by the scattered-field invariant: in vacuum the box carries the full incident
wave while the field outside it stays at float32 round-off, for each propagation axis
and polarization angle.  A sign or plane mistake in the face corrections leaks the wave
out of the box; the last test plants two such mistakes and checks that they do."""
import numpy as np
import pytest

from oracle import oracle as O

N, NPML, LO, HI = 40, 6, 12, 27


def _run(axis, angle, steps=100, mutate=None):
    o = O.OracleSim(N, N, N, 0, NPML, np.float32)
    t = O.OracleTfsf((LO,) * 3, (HI,) * 3, ("fb", "fb", "fb"), 2 * np.pi * 0.5 / 12.0, axis, angle)
    if mutate:
        mutate(t)
    o.tfsf.append(t)
    o.run(steps)
    inside = max(np.abs(f[LO + 1:HI, LO + 1:HI, LO + 1:HI]).max() for f in o.ehs)
    m = np.zeros((N, N, N), bool)
    m[NPML + 1:N - NPML - 1, NPML + 1:N - NPML - 1, NPML + 1:N - NPML - 1] = True
    m[LO - 2:HI + 3, LO - 2:HI + 3, LO - 2:HI + 3] = False
    outside = max(np.abs(f[m]).max() for f in o.ehs)
    return inside, outside


@pytest.mark.parametrize("axis", ["x", "y", "z"])
@pytest.mark.parametrize("angle", [0.0, 0.7])
def test_oracle_tfsf_scattered_field_stays_quiet(axis, angle):
    inside, outside = _run(axis, angle)
    assert inside > 0.5
    assert outside < 1e-5 * inside, (inside, outside)


def test_oracle_tfsf_planted_mistakes_leak():
    def flip_sign(t):
        (geom, (ce1, ce2, ch1, ch2, *sl), nax) = t.pairs[0]
        t.pairs[0] = (geom, (ce1, ce2, -ch1, ch2, *sl), nax)

    def wrong_plane(t):
        (comps, se1, se2, sh1, sh2), coefs, nax = t.pairs[0]
        se1 = (se1[0] + 1,) + tuple(se1[1:])  # E correction one plane inside the box
        t.pairs[0] = ((comps, se1, se2, sh1, sh2), coefs, nax)

    for mutate in (flip_sign, wrong_plane):
        inside, outside = _run("x", 0.0, 60, mutate)
        assert outside > 1e-2 * inside, (mutate.__name__, inside, outside)
