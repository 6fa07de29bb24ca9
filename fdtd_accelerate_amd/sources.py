"""Time functions of the benchmark sources."""
from __future__ import annotations

import math


def gaussian_pulse(t0=100.0, width=30.0, period=20.0):
    """Gaussian-modulated sinusoid in step units (``KB/source/source.py:247-249`` with
    wfreq*dt = 2*pi/period):  exp(-0.5*((t-t0)/width)^2) * cos(2*pi*(t-t0)/period),
    evaluated in float64 (the caller casts to float32, as ``K7/cpu/incident_direct.py:76``)."""

    def tfunc(t):
        x = (t - t0)
        return math.exp(-0.5 * (x / width) ** 2) * math.cos(2.0 * math.pi * x / period)

    return tfunc
