"""Blocks past the 32-bit limits of the addressing.

Population allocations of 4 GiB and more are addressed with 64-bit pointers
(smaller ones through a buffer descriptor with 32-bit offsets), and blocks of
113 M D3Q19 cells and more hold over 2^31 populations, past any 32-bit element
offset. Both are checked by comparing a large block with a small one that
must evolve identically, bit for bit:

  * cavities stacked along z, each closed by its own void layers, evolve
    exactly like one cavity of the same size in a block of its own (every
    cell update reads only its neighbours and flags);
  * a duct periodic along z and driven by a body force along z stays
    z-invariant: every plane of a block of 27,000 planes equals the plane of
    a block of 4 planes, which exercises the periodic wrap across the whole
    allocation.
"""
import numpy as np
import pytest

from freelb_amd import BlockLattice, HalfwayBB, Key_TypePair, abi
from freelb_amd.cases import FLUID, LID, VOID, WALL, cavity_flags

pytestmark = pytest.mark.gpu

ELEM = {"f32": 4, "f16s": 2}


def _cavity(dims, flags, storage, u_lid=0.1, tau=0.6):
    lat = BlockLattice(19, dims, omega=1.0 / tau, storage=storage)
    try:
        lat.upload_flags(flags)
        lat.set_tasks([Key_TypePair(FLUID | WALL | LID, abi.DYN_BGK)])
        lat.set_halfway_bc([HalfwayBB(WALL, VOID, False), HalfwayBB(LID, VOID, True)])
        N = flags.size
        lat.set_field_rho(np.ones(N))
        u = np.zeros((3, N))
        u[0, flags == LID] = u_lid
        lat.set_field_u(u)
        del u
        lat.init_uniform(1.0)
    except Exception:
        lat.close()
        raise
    return lat


@pytest.mark.parametrize("storage,copies", [("f32", 20), ("f16s", 38)])
def test_stacked_cavities_match_one_cavity(storage, copies):
    n, nz = 254, 46  # one cavity: 256 x 256 x 48 slots with its void shell
    one = cavity_flags(n, n, nz)
    dims1 = (n + 2, n + 2, nz + 2)
    dims = (n + 2, n + 2, copies * (nz + 2))
    big_flags = np.tile(one, copies)
    N = big_flags.size
    print(f"stacked cavities {storage}: {copies} x {dims1} = {N} cells", flush=True)
    assert 19 * N * ELEM[storage] > 2 ** 32
    if storage == "f16s":
        assert 19 * N > 2 ** 31
    steps = 30
    small = _cavity(dims1, one, storage)
    try:
        small.step(steps)
        m1 = small.moments()
    finally:
        small.close()
    big = _cavity(dims, big_flags, storage)
    try:
        big.step(steps)
        m = big.moments()
    finally:
        big.close()
    assert np.isfinite(m1).all() and np.abs(m1[1]).max() > 1e-3
    m = m.reshape(4, copies, one.size)
    for k in range(copies):
        assert np.array_equal(m[:, k].view(np.uint32), m1.view(np.uint32)), f"cavity {k}"


def _duct(nz, storage, g=1e-5, tau=0.7):
    """64 x 64 duct cells between halfway walls (void shell in x and y),
    periodic along z with nz planes, body force g along z."""
    f = np.full((nz, 66, 66), VOID, dtype=np.uint8)
    f[:, 1:-1, 1:-1] = FLUID
    lat = BlockLattice(19, (66, 66, nz), omega=1.0 / tau, storage=storage, periodic=(False, False, True))
    try:
        lat.upload_flags(f.reshape(-1))
        lat.set_tasks([Key_TypePair(FLUID, abi.DYN_BGK_FORCE)])
        lat.set_force([0.0, 0.0, g])
        lat.set_halfway_bc([HalfwayBB(FLUID, VOID, False)])
        lat.init_uniform(1.0)
    except Exception:
        lat.close()
        raise
    return lat


def test_periodic_duct_past_2g_populations():
    nz, steps = 27000, 40
    N = 66 * 66 * nz
    print(f"periodic duct f16s: {N} cells", flush=True)
    assert 19 * N > 2 ** 31
    small = _duct(4, "f16s")
    try:
        small.step(steps)
        m1 = small.moments()
    finally:
        small.close()
    big = _duct(nz, "f16s")
    try:
        big.step(steps)
        m = big.moments()
    finally:
        big.close()
    plane1 = m1.reshape(4, 4, 66 * 66)
    for k in range(1, 4):
        assert np.array_equal(plane1[:, k], plane1[:, 0])
    assert plane1[3].max() > 0.0  # the flow along z has started
    m = m.reshape(4, nz, 66 * 66)
    ref = plane1[:, :1].view(np.uint32)
    for k0 in range(0, nz, 3000):
        blk = m[:, k0:k0 + 3000].view(np.uint32)
        assert np.array_equal(blk, np.broadcast_to(ref, blk.shape)), f"planes from {k0}"
