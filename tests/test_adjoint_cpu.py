"""The FP64 CPU model behind the adjoint checks, pinned on its own: its
autograd gradient against its own central differences, exact mass
conservation, exact momentum gain under the Guo force, and the rest state."""
import numpy as np
import pytest
import torch

from adjoint_model import Model
from freelb_amd.cases import solid_fraction


def _case(periodic=(True, False, False), alpha=0.3, g=1e-3, dims=(7, 5, 4), solid=None):
    m = Model(dims, omega=1 / 0.7, force=(g, 2e-4, 0.0), alpha=alpha, periodic=periodic, solid=solid)
    rng = np.random.default_rng(3)
    N = m.N
    f0 = m.equilibrium(1 + 0.01 * rng.standard_normal(N), 0.02 * rng.standard_normal((3, N)))
    s = torch.tensor(rng.uniform(0, 1, N), dtype=torch.float64)
    return m, f0, s


def test_autograd_matches_central_differences():
    m, f0, s = _case()
    s.requires_grad_(True)
    J, _ = m.run(f0, s, 9)
    (g,) = torch.autograd.grad(J, s)
    h = 1e-4
    for cell in (0, 17, 50, m.N - 1):
        sp, sm = s.detach().clone(), s.detach().clone()
        sp[cell] += h
        sm[cell] -= h
        fd = (m.run(f0, sp, 9)[0] - m.run(f0, sm, 9)[0]).item() / (2 * h)
        assert abs(fd - g[cell].item()) <= 1e-6 * abs(fd) + 1e-14


@pytest.mark.parametrize("periodic", [(True, False, False), (False, False, False)])
def test_mass_is_conserved(periodic):
    solid = np.zeros(7 * 5 * 4, dtype=bool)
    solid[[12, 40, 41]] = True
    m, f0, s = _case(periodic=periodic, solid=solid)
    f = f0
    m0 = f.sum().item()
    for _ in range(11):
        f = m.step(f, s)
    assert abs(f.sum().item() - m0) < 1e-12 * m0
    assert torch.all(f[:, torch.tensor(solid)] == 0)


def test_guo_force_adds_exactly_G_per_cell_and_step():
    m, f0, s = _case(periodic=(True, True, True), alpha=0.0)
    j0 = (m.C.T @ f0).sum(1)
    f = f0
    for _ in range(5):
        f = m.step(f, s)
    j5 = (m.C.T @ f).sum(1)
    np.testing.assert_allclose((j5 - j0).numpy(), 5 * m.N * m.G.numpy(), rtol=1e-10, atol=1e-13)


def test_rest_state_is_fixed_without_force():
    m = Model((6, 4, 3), omega=1.3, force=(0, 0, 0), alpha=0.5)
    f0 = m.equilibrium(np.ones(m.N), np.zeros((3, m.N)))
    f = m.step(f0, torch.full((m.N,), 0.4, dtype=torch.float64))
    assert torch.allclose(f, f0, rtol=0, atol=1e-16)


def test_solid_fraction_field():
    s = solid_fraction((24, 12, 10), seed=4)
    assert s.shape == (24 * 12 * 10,)
    assert 0.0 < s.min() and s.max() < 1.0
    assert abs(s.mean() - 0.5) < 0.05
    assert np.array_equal(s, solid_fraction((24, 12, 10), seed=4))
