import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built libmifdtd.so")


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    return np.load(os.path.join(ROOT, "tests", "golden", "kemp_core.npz"))


def golden_cases(g):
    """Yield (meta dict, inputs, ces|None, chs|None, outputs) of the golden file."""
    for c in range(int(g["ncases"])):
        p = f"c{c}_"
        nx, ny, nz, nzp, pad, tmax, uf, ce_on, ch_on = [int(v) for v in g[p + "meta"]]
        meta = dict(nx=nx, ny=ny, nz=nz, nzp=nzp, pad=pad, tmax=tmax, ufunc=("e", "h", "eh")[uf],
                    ce_on=bool(ce_on), ch_on=bool(ch_on), case=c)
        ces = list(g[p + "ce"]) if ce_on else None
        chs = list(g[p + "ch"]) if ch_on else None
        yield meta, list(g[p + "in"]), ces, chs, list(g[p + "out"])
