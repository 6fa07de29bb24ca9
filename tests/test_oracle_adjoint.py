"""Pins of the FP64 reference gradient used to check the adjoint kernels: the float64
torch restatement of the step equals the C oracle, and its autograd gradient equals
central finite differences."""
import numpy as np
import torch

from oracle import oracle as O
from oracle.torch_ref import TorchYee, objective_grad


def test_torch_restatement_equals_c_oracle():
    nx, ny, nz, pad, npml = 20, 18, 22, 2, 4
    rng = np.random.default_rng(0)
    ces = [(0.5 / rng.uniform(1, 4, (nx, ny, nz + pad))) for _ in range(3)]
    o = O.OracleSim(nx, ny, nz, pad, npml, np.float64, ces=ces)
    for a in o.ehs:
        a[:] = rng.standard_normal(a.shape)
        a[:, :, nz:] = 0
    mask = np.zeros((nx, ny, nz + pad))
    mask[6, npml:ny - npml, npml:nz - npml] = 1
    o.add_source("ez", (6, npml, npml), (6, ny - npml - 1, nz - npml - 1), O.pulse)
    m = TorchYee(nx, ny, nz, pad, npml)
    st = m.zeros()
    for k in range(6):
        st[k] = torch.tensor(o.ehs[k])
    ce_t = [torch.tensor(c) for c in ces]
    with torch.no_grad():
        for n in range(1, 7):
            o.step()
            st = m.step(st, ce_t, [(2, O.pulse(n) * torch.tensor(mask))])
    for k in range(6):
        np.testing.assert_array_equal(o.ehs[k], st[k].numpy())


def test_autograd_gradient_equals_central_differences():
    nx, ny, nz, pad, npml = 24, 22, 20, 4, 4
    box = (9, 8, 7, 5, 5, 5)
    m = TorchYee(nx, ny, nz, pad, npml)
    ceb = [torch.full((nx, ny, nz + pad), 0.5, dtype=torch.float64) for _ in range(3)]
    mask = torch.zeros((nx, ny, nz + pad), dtype=torch.float64)
    mask[5, npml:ny - npml, npml:nz - npml] = 1
    tf = lambda t: O.pulse(t, 20.0, 6.0, 10.0)
    src = lambda n: [(2, tf(n) * mask)]
    eps = torch.tensor(np.random.default_rng(1).uniform(1, 4, box[3:]))
    J, g = objective_grad(m, ceb, box, eps, 80, 61, 18, src)
    h = 1e-5
    for v in [(0, 0, 0), (2, 2, 2), (4, 1, 3)]:
        e1, e2 = eps.clone(), eps.clone()
        e1[v] += h
        e2[v] -= h
        fd = (objective_grad(m, ceb, box, e1, 80, 61, 18, src, want_grad=False)[0] -
              objective_grad(m, ceb, box, e2, 80, 61, 18, src, want_grad=False)[0]) / (2 * h)
        assert abs(fd - float(g[v])) <= 1e-7 * abs(float(g[v])) + 1e-12, (v, fd, float(g[v]))


def test_two_level_reverse_sweep_equals_checkpointed():
    """The segment-by-segment reverse sweep (states every `outer` steps) gives the same
    J and gradient as the plain checkpointed backward."""
    nx, ny, nz, pad, npml = 24, 22, 20, 4, 4
    box = (9, 8, 7, 5, 5, 5)
    m = TorchYee(nx, ny, nz, pad, npml)
    ceb = [torch.full((nx, ny, nz + pad), 0.5, dtype=torch.float64) for _ in range(3)]
    mask = torch.zeros((nx, ny, nz + pad), dtype=torch.float64)
    mask[5, npml:ny - npml, npml:nz - npml] = 1
    tf = lambda t: O.pulse(t, 20.0, 6.0, 10.0)
    src = lambda n: [(2, tf(n) * mask)]
    eps = torch.tensor(np.random.default_rng(1).uniform(1, 4, box[3:]))
    J1, g1 = objective_grad(m, ceb, box, eps, 80, 61, 18, src, seg=10)
    J2, g2 = objective_grad(m, ceb, box, eps, 80, 61, 18, src, seg=3, outer=17)
    assert abs(J1 - J2) <= 1e-12 * abs(J1)
    assert float((g1 - g2).abs().max()) <= 1e-12 * float(g1.abs().max())
