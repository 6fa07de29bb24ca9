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


def split_cases():
    """(meta dict, inputs (6 arrays), coefficient arrays (3), outputs of the 3 updated
    components) of tests/golden/kemp_core_split.npz, the reference's split programs."""
    import numpy as np
    g = np.load(os.path.join(ROOT, "tests", "golden", "kemp_core_split.npz"))
    for c in range(int(g["ncases"])):
        m = g[f"c{c}_meta"]
        meta = dict(zip(("nx", "ny", "nz", "nzp", "pad", "idx0", "nmax", "h", "f64"), (int(v) for v in m)))
        yield meta, list(g[f"c{c}_in"]), list(g[f"c{c}_coef"]), list(g[f"c{c}_out"])


def split_expected(meta, ehs, coef):
    """The split form from the oracle's whole-grid update: a vector (4 float / 2 double
    elements) takes the updated values when its first element lies in [idx0, nmax)."""
    import numpy as np
    from oracle import oracle as O
    full = [x.copy() for x in ehs]
    (O.update_h if meta["h"] else O.update_e)(full, coef, meta["pad"])
    vec = 2 if meta["f64"] else 4
    flat = np.arange(ehs[0].size).reshape(ehs[0].shape)
    start = flat // vec * vec
    inside = (start >= meta["idx0"]) & (start < meta["nmax"])
    sl = slice(3, 6) if meta["h"] else slice(0, 3)
    return [np.where(inside, f, x) for f, x in zip(full[sl], ehs[sl])]
