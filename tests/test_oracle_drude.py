"""The oracle's Drude dispersion (auxiliary current, ``KU/dim3/cpu/core/drude_ade.c``
with the coefficients of ``KU/dim3/cpu/matter.py:192-202``), pinned by the plasma
oscillation of a uniform periodic medium.  A uniform E has zero curl, so H stays 0 and
each step is e' = e - CE J, J' = CJA J + CJB e'.  Without damping the uniform E then
obeys the discrete oscillator e(n+2) = (2 - (wp dt)^2) e(n+1) - e(n), whose frequency
per step arccos(1 - (wp dt)^2 / 2) tends to wp dt; with damping the amplitude falls by
sqrt((2 - g dt) / (2 + g dt)) per step.  A wrong CJB factor, the old e in the current
update, or a missing CE breaks these; the last test plants such mistakes."""
import numpy as np
import pytest

from oracle import oracle as O

DT = 0.5


def _series(wp, gamma, steps=160, dtype=np.float64, mutate=None):
    n = 7
    cja, cjb = O.drude_coefficients(wp, gamma, DT, dtype)
    o = O.OracleSim(n, n, n, 0, 0, dtype, pbc="xyz")
    for k in range(3):
        o.ehs[k][...] = 1.0
    d = O.OracleDrude((0, 0, 0), (n - 1,) * 3, [cja] * 3, [cjb] * 3, dtype)
    if mutate:
        mutate(d)
    o.drude.append(d)
    e = []
    for _ in range(steps):
        o.step()
        e.append(o.ehs[2][3, 3, 3])
    assert max(np.abs(o.ehs[k]).max() for k in (3, 4, 5)) == 0.0
    assert np.ptp(o.ehs[0]) == 0.0
    return np.array(e)


@pytest.mark.parametrize("wp", [0.1, 0.4, 1.2])
def test_oracle_drude_plasma_oscillation(wp):
    e = _series(wp, 0.0)
    c = 2.0 - (wp * DT) ** 2
    assert np.abs(e[2:] - (c * e[1:-1] - e[:-2])).max() < 1e-12
    # the frequency per step approaches wp dt as wp dt -> 0 (second order)
    theta = np.arccos(1 - (wp * DT) ** 2 / 2)
    assert abs(theta - wp * DT) < (wp * DT) ** 3 / 20


@pytest.mark.parametrize("wp,gamma", [(0.5, 0.1), (1.0, 0.4)])
def test_oracle_drude_damping(wp, gamma):
    e = _series(wp, gamma, 400)
    r2 = (2 - gamma * DT) / (2 + gamma * DT)
    # the 2x2 step matrix has determinant r2: e(n+2) - tr e(n+1) + r2 e(n) = 0 for one tr
    tr = (e[2:] + r2 * e[:-2]) / e[1:-1]
    ok = np.abs(e[1:-1]) > 1e-6 * np.abs(e).max()
    assert np.ptp(tr[ok]) < 1e-8, np.ptp(tr[ok])
    assert abs(tr[ok].mean() - (1 + r2 - 0.5 * 4 * (DT * wp) ** 2 / (2 + gamma * DT))) < 1e-10


def test_oracle_drude_planted_mistakes_fail():
    def old_e(d):
        def upd(ehs, ces):
            for c in range(3):
                e = ehs[c][d.sl].copy()
                ehs[c][d.sl] = e - 0.5 * d.J[c]
                d.J[c] = d.cja[c] * d.J[c] + d.cjb[c] * e
        d.update_e = upd

    def half_cjb(d):
        d.cjb = [c / 2 for c in d.cjb]

    wp = 0.4
    c = 2.0 - (wp * DT) ** 2
    for mutate in (old_e, half_cjb):
        e = _series(wp, 0.0, mutate=mutate)
        assert np.abs(e[2:] - (c * e[1:-1] - e[:-2])).max() > 1e-4, mutate.__name__
