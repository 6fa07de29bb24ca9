"""Full-size parity of the fused engine on the 1024^3 throughput workload (200-sphere
dielectric, CPML 12 on all faces, Ez point source and probe, CE held as uint8 material
indices) against the in-place engine reading the float CE arrays: every field, every
psi value and every probe sample equal after odd and even step counts (values compared
with -0 == +0, the fused kernel's masked lanes may store +0).

This is the shape the throughput figure is measured on, and the only one the small
parity tests cannot reach: field sets carved from one staggered allocation (arrays of
1 GiB and more), four 1 KiB segments per row, the interior box on 128-byte z lines and
the full block grid of the sweep and the shell passes."""
import gc

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

NAMES = ("ex", "ey", "ez", "hx", "hy", "hz")


def _build(engine, ce_indices):
    import bench

    class Args:
        n, npml, spheres = 1024, 12, 200
    a = Args()
    a.engine = engine
    a.no_ce_lut = not ce_indices
    return bench.build_sim(a)


def test_bench_workload_fused_equals_inplace():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from fdtd_accelerate_amd import _lib
    _lib.load()
    gc.collect()
    torch.cuda.empty_cache()
    free, _ = torch.cuda.mem_get_info()
    if free < 160e9:
        pytest.skip(f"needs about 140 GB of free device memory, {free / 1e9:.0f} GB free")
    ref = _build("kemp", False)
    fus = _build("fused", True)
    try:
        assert ref.fields.ce_mat is None and fus.fields.ce_mat is not None
        for steps in (3, 2):
            a = ref.run(steps)
            b = fus.run(steps)
            assert np.array_equal(a + np.float32(0), b + np.float32(0)), (steps, a, b)
            for name in NAMES:
                assert torch.equal(ref.field(name), fus.field(name)), (steps, name)
            for key in ref.cpml.psiE:
                for c in range(2):
                    assert torch.equal(ref.cpml.psiE[key][c], fus.cpml.psiE[key][c]), ("psiE", steps, key, c)
                    assert torch.equal(ref.cpml.psiH[key][c], fus.cpml.psiH[key][c]), ("psiH", steps, key, c)
        arrays = fus.fields.ehs + fus._fused.alt
        base = arrays[0].untyped_storage().data_ptr()
        assert all(t.untyped_storage().data_ptr() == base for t in arrays)
        assert float(torch.abs(fus.field("ez")).max()) > 0.0
    finally:
        del ref, fus
        gc.collect()
        torch.cuda.empty_cache()
