"""The test twins themselves: the periodic-x twin against a tiled domain.

A domain periodic in x evolves as the infinite periodic tiling of its
state. Three copies of a channel side by side, with the tiling's ends closed
by ordinary walls, keep the middle copy equal to the periodic solution for as
many steps as information needs to cross one copy (one cell per step). The
twin with its mirrored virtual columns must reproduce that middle copy bit
for bit.
"""
import numpy as np

from oracle import oracle as O
from tests.twins import PeriodicXTwin, cavity_twin, interior


def _channel_flags(nx, ny, nz):
    f = np.full((nz + 2, ny + 2, nx), O.VOID, dtype=np.uint8)
    f[1:-1, 1:-1, :] = O.AABB
    return f


def test_periodic_twin_matches_tiled_domain():
    nx, ny, nz = 11, 6, 5
    f = _channel_flags(nx, ny, nz)
    rng = np.random.default_rng(3)
    N = f.size
    rhou = np.zeros((4, N))
    rhou[0] = 1 + 0.01 * rng.standard_normal(N)
    rhou[1:] = 0.02 * rng.standard_normal((3, N))
    tasks = [(O.AABB, O.DYN_BGK)]
    bcs = [(O.AABB, O.VOID, O.HWBC_NORMAL)]
    twin = PeriodicXTwin(19, (nx, ny + 2, nz + 2), f.reshape(-1), tasks, bcs, omega=1 / 0.6)
    p0 = twin.lat.init_feq(_widen4(rhou, nx, ny + 2, nz + 2))
    p0 = twin.narrow(p0)
    # tiled: three copies along x plus a wall column at each end
    tf = np.concatenate([f[:, :, :1] * 0 + O.VOID, f, f, f, f[:, :, :1] * 0 + O.VOID], axis=2)
    tiled = O.Lattice(19, (3 * nx + 2, ny + 2, nz + 2), tf.reshape(-1), tasks, bcs, omega=1 / 0.6)
    p4 = p0.reshape(19, nz + 2, ny + 2, nx)
    zero = np.zeros_like(p4[..., :1])
    tp = np.ascontiguousarray(np.concatenate([zero, p4, p4, p4, zero], axis=3).reshape(19, -1))
    steps = nx - 1
    tiled.run(tp, steps)
    mid = tp.reshape(19, nz + 2, ny + 2, 3 * nx + 2)[..., 1 + nx:1 + 2 * nx].reshape(19, -1)
    wp = twin.widen(p0)
    twin.run(wp, steps)
    got = twin.narrow(wp)
    live = interior((nx, ny + 2, nz + 2), periodic=(True, False, False))
    assert np.array_equal(got[:, live], mid[:, live])
    # and the state did move
    assert np.abs(got[:, live] - p0[:, live]).max() > 1e-4


def _widen4(rhou, nx, ny, nz):
    a = rhou.reshape(4, nz, ny, nx)
    return np.ascontiguousarray(np.concatenate([a[..., -1:], a, a[..., :1]], axis=3).reshape(4, -1))


def test_cavity_twin_flags_match_the_oracle_cavity():
    """The twin's cavity flags are the oracle's FreeLB cavity flags."""
    for shape in [(8, 7, 6), (20, 9, 13)]:
        t = cavity_twin(shape)
        assert np.array_equal(t.flags, O.cavity_flags(*shape))
