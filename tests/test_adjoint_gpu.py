"""The hand-written GPU adjoint (through the C ABI) against autodiff of the
FP64 CPU model and against central finite differences in FP64; checkpointed
recompute against keeping every state; FP32 against FP64; and the channel of
192 x 96 x 96 cells, T = 800, checkpoints every 28 steps, against FP64 finite
differences (h = 1e-4) at 32 seeded cells."""
import numpy as np
import pytest
import torch

from adjoint_model import Model
from adjoint_torch import AdjointObjective
from freelb_amd import Adjoint, AdjointBatch
from freelb_amd.cases import channel_adjoint

pytestmark = pytest.mark.gpu

DIMS = (10, 6, 5)
TAU, G, ALPHA = 0.7, (1e-3, 2e-4, 0.0), 0.3
PER = (True, False, False)


def _setup(storage="f64"):
    N = int(np.prod(DIMS))
    rng = np.random.default_rng(21)
    solid = np.zeros(N, dtype=bool)
    solid[[33, 34, 140, 201]] = True
    rho = 1 + 0.01 * rng.standard_normal(N)
    u = 0.02 * rng.standard_normal((3, N))
    s = rng.uniform(0, 1, N)
    adj = Adjoint(DIMS, omega=1 / TAU, force=G, brinkman_alpha=ALPHA, periodic=PER, storage=storage)
    adj.set_solid(solid)
    adj.set_fraction(s)
    adj.init_equilibrium(rho, u)
    model = Model(DIMS, omega=1 / TAU, force=G, alpha=ALPHA, periodic=PER, solid=solid)
    f0 = model.equilibrium(rho, u)
    return adj, model, f0, s, solid


def test_forward_matches_model_fp64():
    adj, model, f0, s, solid = _setup()
    with adj:
        J = adj.forward(15, -1)
        Jm, fm = model.run(f0, torch.tensor(s), 15)
        assert abs(J - Jm.item()) <= 1e-12 * abs(Jm.item())
        np.testing.assert_allclose(adj.state(), fm.numpy(), rtol=0, atol=1e-14)


@pytest.mark.parametrize("dims,per", [
    (DIMS, (True, False, False)), (DIMS, (False, False, False)), (DIMS, (False, True, True)),
    # extents of 1: the cell index divides by 1, and both faces of an axis
    # fall on one cell (or a periodic axis wraps onto the cell itself)
    ((1, 6, 5), (True, False, False)), ((8, 1, 5), (False, False, True)), ((6, 4, 1), (False, False, False))])
def test_solid_free_box_matches_model_fp64(dims, per):
    """Without solid cells the steps take each cell's wall links from its
    place (the faces of the non-periodic axes) instead of loading link words:
    objective, final state and gradient against the FP64 model."""
    N = int(np.prod(dims))
    rng = np.random.default_rng(23)
    solid = np.zeros(N, dtype=bool)
    rho = 1 + 0.01 * rng.standard_normal(N)
    u = 0.02 * rng.standard_normal((3, N))
    s = rng.uniform(0, 1, N)
    model = Model(dims, omega=1 / TAU, force=G, alpha=ALPHA, periodic=per, solid=solid)
    f0 = model.equilibrium(rho, u)
    with Adjoint(dims, omega=1 / TAU, force=G, brinkman_alpha=ALPHA, periodic=per, storage="f64") as adj:
        adj.set_fraction(s)
        adj.init_equilibrium(rho, u)
        J = adj.forward(15, 4)
        state = adj.state()
        g = adj.backward()
    st = torch.tensor(s, requires_grad=True)
    Jm, fm = model.run(f0, st, 15)
    (gm,) = torch.autograd.grad(Jm, st)
    assert abs(J - Jm.item()) <= 1e-12 * abs(Jm.item())
    np.testing.assert_allclose(state, fm.detach().numpy(), rtol=0, atol=1e-14)
    np.testing.assert_allclose(g, gm.numpy(), rtol=0, atol=1e-9 * np.abs(gm.numpy()).max())


@pytest.mark.parametrize("K", [0, 1, 4, 15, 100])
def test_gradient_matches_autodiff_fp64(K):
    adj, model, f0, s, solid = _setup()
    with adj:
        adj.forward(15, K)
        g = adj.backward()
    st = torch.tensor(s, requires_grad=True)
    Jm, _ = model.run(f0, st, 15)
    (gm,) = torch.autograd.grad(Jm, st)
    gm = gm.numpy()
    assert np.all(g[solid] == 0)
    np.testing.assert_allclose(g, gm, rtol=0, atol=1e-9 * np.abs(gm).max())


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_checkpoint_modes_agree_bitwise(storage):
    """Every checkpoint interval gives the same objective, final state and
    gradient bit for bit: moment records written by the forward pass (K 0)
    equal those of the recomputed segments (K 5, 7; K 40 > T is one segment)."""
    adj, *_ = _setup(storage)
    out = []
    with adj:
        for K in (0, 5, 7, 40):
            J = adj.forward(23, K)
            out.append((J, adj.state(), adj.backward()))
    J0, s0, g0 = out[0]
    for J, st, g in out[1:]:
        assert J == J0 and np.array_equal(st, s0) and np.array_equal(g, g0)


def test_gradient_against_fd_fp64():
    adj, model, f0, s, solid = _setup()
    with adj:
        adj.forward(15, 4)
        g = adj.backward()
        h = 1e-4
        for cell in (0, 57, 123, 299):
            adj.set_fraction_at(cell, s[cell] + h)
            jp = adj.forward(15, -1)
            adj.set_fraction_at(cell, s[cell] - h)
            jm = adj.forward(15, -1)
            adj.set_fraction_at(cell, s[cell])
            fd = (jp - jm) / (2 * h)
            assert abs(g[cell] - fd) <= 1e-6 * abs(fd) + 1e-14, (cell, g[cell], fd)


def test_fp32_adjoint_close_to_fp64():
    a64, *_ = _setup("f64")
    a32, *_ = _setup("f32")
    with a64, a32:
        a64.forward(15, 4)
        a32.forward(15, 4)
        g64, g32 = a64.backward(), a32.backward()
    assert np.abs(g32 - g64).max() <= 1e-4 * np.abs(g64).max()


def test_autograd_function_wrapper():
    adj, model, f0, s, solid = _setup()
    with adj:
        st = torch.tensor(s, requires_grad=True)
        J = AdjointObjective.apply(st, adj, 15, 4)
        (2.0 * J).backward()
        g = adj.backward()
    np.testing.assert_allclose(st.grad.numpy(), 2.0 * g, rtol=0, atol=0)


def test_channel_gradient_against_fd():
    """192 x 96 x 96 channel, periodic x, walls in y and z, g = 1e-6, tau 0.6,
    Brinkman s from the seeded smoothed field, T = 800: FP32 adjoint with
    checkpoints every 28 steps against FP64 central differences."""
    T, K, h = 800, 28, 1e-4
    adj32, s = channel_adjoint(storage="f32")
    with adj32:
        adj32.forward(T, K)
        g = adj32.backward()
        fwd_ms, bwd_ms = adj32.timing()
    N = s.size
    cells = np.random.default_rng(99).choice(N, 32, replace=False)
    adj64, s64 = channel_adjoint(storage="f64")
    with adj64:
        fd = np.empty(32)
        for k, c in enumerate(cells):
            adj64.set_fraction_at(int(c), s64[c] + h)
            jp = adj64.forward(T, -1)
            adj64.set_fraction_at(int(c), s64[c] - h)
            jm = adj64.forward(T, -1)
            adj64.set_fraction_at(int(c), s64[c])
            fd[k] = (jp - jm) / (2 * h)
    rel = np.abs(g[cells] - fd) / np.abs(fd)
    print(f"channel: forward {fwd_ms:.1f} ms, backward {bwd_ms:.1f} ms, max rel {rel.max():.3e}, "
          f"median rel {np.median(rel):.3e}, |fd| in [{np.abs(fd).min():.3e}, {np.abs(fd).max():.3e}]")
    assert rel.max() <= 1e-3


def _mixed_members(storage):
    """Four boxes of different sizes and boundary setups (with and without
    solid cells)."""
    specs = [((10, 6, 5), (True, False, False), [33, 34, 140, 201]),
             ((7, 9, 4), (False, True, False), [5, 77, 150]),
             ((13, 5, 6), (True, True, True), []),
             ((9, 7, 6), (False, False, True), [])]
    out = []
    for k, (dims, per, sol) in enumerate(specs):
        N = int(np.prod(dims))
        rng = np.random.default_rng(40 + k)
        solid = np.zeros(N, dtype=bool)
        solid[sol] = True
        a = Adjoint(dims, omega=1 / (0.6 + 0.1 * k), force=(1e-3, -2e-4 * k, 1e-4), brinkman_alpha=0.2 + 0.1 * k,
                    outlet_x=dims[0] - 1 - k, periodic=per, storage=storage)
        a.set_solid(solid)
        a.set_fraction(rng.uniform(0, 1, N))
        a.init_equilibrium(1 + 0.01 * rng.standard_normal(N), 0.02 * rng.standard_normal((3, N)))
        out.append(a)
    return out


@pytest.mark.parametrize("storage", ["f32", "f64"])
@pytest.mark.parametrize("K", [0, 4, 6])
def test_batch_matches_members_bitwise(storage, K):
    """The batched forward and adjoint give each member exactly what its own
    run gives: objective, final state and gradient, bit for bit."""
    T = 17
    solo = _mixed_members(storage)
    mem = _mixed_members(storage)
    try:
        J1, S1, G1 = [], [], []
        for a in solo:
            J1.append(a.forward(T, K))
            S1.append(a.state())
            G1.append(a.backward())
        with AdjointBatch(mem) as b:
            J2 = b.forward(T, K)
            S2 = [a.state() for a in mem]
            G2 = b.backward()
            # a member's own backward after the batch forward run
            G3 = mem[1].backward()
        assert np.array_equal(np.array(J1), J2)
        for s1, s2 in zip(S1, S2):
            assert np.array_equal(s1, s2)
        for g1, g2 in zip(G1, G2):
            assert np.array_equal(g1, g2)
        assert np.array_equal(G3, G1[1])
    finally:
        for a in solo + mem:
            a.close()


def test_batch_forward_only_and_rerun():
    mem = _mixed_members("f32")
    try:
        with AdjointBatch(mem) as b:
            J = b.forward(9, -1)
            with pytest.raises(Exception):
                b.backward()
            assert np.array_equal(J, b.forward(9, -1))
            b.forward(12, 5)
            g = b.backward()
            assert len(g) == len(mem) and all(x.size == m.N for x, m in zip(g, mem))
    finally:
        for a in mem:
            a.close()


def test_batch_rejects_bad_members():
    a32 = _mixed_members("f32")
    a64 = _mixed_members("f64")
    try:
        with pytest.raises(Exception):
            AdjointBatch([a32[0], a32[0]])
        with pytest.raises(Exception):
            AdjointBatch([a32[0], a64[1]])
    finally:
        for a in a32 + a64:
            a.close()


def test_channel_batch_runs_and_matches_one_member():
    """Eight channels of seeded extents in [24,12,12]..[40,20,20]: the batch
    gradient of one member equals that member's own run."""
    from freelb_amd.cases import channel_batch, channel_dims
    dims = channel_dims(8, lo=(24, 12, 12), hi=(40, 20, 20), seed=3)
    b, mem, fr = channel_batch(dims, seed=3)
    try:
        J = b.forward(40, 7)
        G = b.backward()
        assert np.all(np.isfinite(J)) and np.all(J > 0)
        solo, _ = channel_adjoint(dims[5], seed=1000 * 3 + 5)
        with solo:
            assert solo.forward(40, 7) == J[5]
            assert np.array_equal(solo.backward(), G[5])
    finally:
        b.close()
        for a in mem:
            a.close()
