"""Same-level block halo: a domain cut into blocks whose one-cell overlap
layers mirror their neighbours, exchanged after every step, evolves exactly
as the same domain in one block (bit for bit), with per-block steps, in a
batch, replayed from a hipGraph, and across a periodic domain axis. In FP32
the blocks are also held against the oracle stepping the whole domain (for
the periodic axis, its periodic twin, tests/twins.py)."""
import numpy as np
import pytest

from freelb_amd import (BlockHalo, BlockLattice, BlockLatticeManager, HalfwayBB, Key_TypePair, abi)
from gpu_case import gpu_lattice
from oracle import oracle as O
from tests.twins import PeriodicXTwin

pytestmark = pytest.mark.gpu


def _intervals(n, cuts):
    edges = [1] + list(cuts) + [n + 1]
    return [(edges[k], edges[k + 1]) for k in range(len(edges) - 1)]


def _sub(arr, dims, box, comps=1):
    """Slice a dense [comps][N] global array (x fastest) to a local box."""
    nx, ny, nz = dims
    (x0, x1), (y0, y1), (z0, z1) = box
    a = np.asarray(arr).reshape(comps, nz, ny, nx)
    return np.ascontiguousarray(a[:, z0:z1, y0:y1, x0:x1].reshape(comps, -1))


def _split_cavity(case, pops, cuts, storage="f32"):
    dims = (case.nx, case.ny, case.nz)
    blocks, origins, boxes = [], [], []
    for (x0, x1) in _intervals(case.nx - 2, cuts[0]):
        for (y0, y1) in _intervals(case.ny - 2, cuts[1]):
            for (z0, z1) in _intervals(case.nz - 2, cuts[2]):
                box = ((x0 - 1, x1 + 1), (y0 - 1, y1 + 1), (z0 - 1, z1 + 1))
                ldims = tuple(b - a for a, b in box)
                lat = BlockLattice(19, ldims, omega=case.omega, storage=storage)
                lat.set_ghost_overlap(True)
                lat.upload_flags(_sub(case.flags, dims, box)[0])
                lat.set_tasks([Key_TypePair(t[0], t[1], float(t[2]) if len(t) > 2 else 0.0) for t in case.tasks])
                lat.set_halfway_bc([HalfwayBB(b[0], b[1], b[2] == 1) for b in case.bcs])
                lat.set_field_rho(_sub(case.rho_field, dims, box)[0])
                lat.set_field_u(_sub(case.u_field, dims, box, 3).reshape(-1))
                lat.upload_pops(_sub(pops, dims, box, 19))
                blocks.append(lat)
                origins.append((x0 - 1, y0 - 1, z0 - 1))
                boxes.append(box)
    return blocks, origins, boxes


def _compare(ref, blocks, boxes, dims, sel):
    """Each block's updated cells (sel, overlap cut) against the whole-domain
    populations ref."""
    for lat, box in zip(blocks, boxes):
        mine = lat.pops().reshape(19, *tuple(b - a for a, b in box)[::-1])
        want = _sub(ref, dims, box, 19).reshape(mine.shape)
        m = _sub(sel.astype(np.uint8), dims, box)[0].reshape(mine.shape[1:]).astype(bool)
        inner = np.zeros_like(m)
        inner[1:-1, 1:-1, 1:-1] = True
        m &= inner
        assert np.array_equal(mine[:, m], want[:, m]), f"block at {box}"


@pytest.mark.parametrize("mode", ["hw", "fw", "smag"])
@pytest.mark.parametrize("storage", ["f32", "f16s"])
def test_split_cavity_matches_single_block(mode, storage):
    case, p = O.cavity_case(mode, 23, 17, 13)
    dims = (case.nx, case.ny, case.nz)
    glob = gpu_lattice(case, storage, p)
    blocks, origins, boxes = _split_cavity(case, p, ([9], [5, 12], [7]), storage)
    halo = BlockHalo(blocks, origins)
    assert halo.info()["items"] > 0
    sel = case.flags != O.VOID if mode == "hw" else case.flags == O.AABB
    ref = p.copy() if storage == "f32" else None
    try:
        for n in (1, 1, 3, 4):
            glob.step(n)
            for _ in range(n):
                for b in blocks:
                    b.step(1)
                halo.exchange()
            _compare(glob.pops(), blocks, boxes, dims, sel)
            if ref is not None:
                case.run(ref, n)
                _compare(ref, blocks, boxes, dims, sel)
    finally:
        halo.close()
        for b in blocks:
            b.close()
        glob.close()


def test_split_cavity_in_batch_and_graph():
    case, p = O.cavity_case("hw", 30, 21, 18)
    dims = (case.nx, case.ny, case.nz)
    glob = gpu_lattice(case, "f32", p)
    blocks, origins, boxes = _split_cavity(case, p, ([11, 20], [10], [6, 12]))
    halo = BlockHalo(blocks, origins)
    sel = case.flags != O.VOID
    ref = p.copy()
    try:
        with BlockLatticeManager(blocks, halo=halo) as m:
            m.step(5)
            glob.step(5)
            m.synchronize()
            _compare(glob.pops(), blocks, boxes, dims, sel)
            case.run(ref, 5)
            _compare(ref, blocks, boxes, dims, sel)
            m.capture(6)
            for _ in range(3):
                m.replay()
                glob.step(6)
                m.synchronize()
                _compare(glob.pops(), blocks, boxes, dims, sel)
                case.run(ref, 6)
                _compare(ref, blocks, boxes, dims, sel)
    finally:
        halo.close()
        for b in blocks:
            b.close()
        glob.close()


def test_periodic_axis_across_blocks():
    """A channel periodic in x, walls in y and z: three blocks along x with
    the wrap carried by the halo against one block with a periodic x axis."""
    nx, ny, nz = 40, 9, 7
    gdims = (nx, ny + 2, nz + 2)
    f = np.full((nz + 2, ny + 2, nx), 1, dtype=np.uint8)
    f[1:-1, 1:-1, :] = 2
    flags = f.reshape(-1)
    N = flags.size
    rng = np.random.default_rng(4)
    u0 = np.zeros((3, N))
    u0[:, flags == 2] = rng.uniform(-0.03, 0.03, size=(3, int((flags == 2).sum())))

    def setup(lat, fl):
        lat.upload_flags(fl)
        lat.set_tasks([Key_TypePair(2, abi.DYN_BGK)])
        lat.set_halfway_bc([HalfwayBB(2, 1, False)])

    glob = BlockLattice(19, gdims, omega=1 / 0.7, periodic=(True, False, False))
    setup(glob, flags)
    glob.init_equilibrium(np.ones(N), u0)
    p0 = glob.pops()
    blocks, origins, boxes = [], [], []
    for (a, b) in ((0, 13), (13, 27), (27, 40)):
        xs = np.arange(a - 1, b + 1) % nx
        ldims = (b - a + 2, ny + 2, nz + 2)
        lat = BlockLattice(19, ldims, omega=1 / 0.7)
        lat.set_ghost_overlap(True)
        setup(lat, np.ascontiguousarray(f[:, :, xs]).reshape(-1))
        pl = p0.reshape(19, nz + 2, ny + 2, nx)[:, :, :, xs]
        lat.upload_pops(np.ascontiguousarray(pl.reshape(19, -1)))
        blocks.append(lat)
        origins.append((a - 1, 0, 0))
        boxes.append(xs)
    halo = BlockHalo(blocks, origins, periodic=(True, False, False), extent=(nx, 0, 0))
    twin = PeriodicXTwin(19, gdims, flags, [(2, O.DYN_BGK)], [(2, 1, O.HWBC_NORMAL)], omega=1 / 0.7)
    wide = twin.widen(p0)
    try:
        for n in (1, 2, 5):
            glob.step(n)
            for _ in range(n):
                for lat in blocks:
                    lat.step(1)
                halo.exchange()
            twin.run(wide, n)
            for ref in (glob.pops(), twin.narrow(wide)):
                ref = ref.reshape(19, nz + 2, ny + 2, nx)
                for lat, xs in zip(blocks, boxes):
                    mine = lat.pops().reshape(19, nz + 2, ny + 2, len(xs))
                    inner = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
                    assert np.array_equal(mine[inner], ref[:, 1:-1, 1:-1, xs[1:-1]])
    finally:
        halo.close()
        for lat in blocks:
            lat.close()
        glob.close()


def test_halo_rejects_bad_layouts():
    case, p = O.cavity_case("hw", 12, 10, 8)
    blocks, origins, boxes = _split_cavity(case, p, ([6], [], []))
    try:
        bad = list(origins)
        bad[1] = (bad[1][0] - 2, bad[1][1], bad[1][2])  # the two interiors overlap
        with pytest.raises(Exception):
            BlockHalo(blocks, bad)
        blocks[0].step(1)  # parity mismatch at exchange
        halo = BlockHalo(blocks, origins)
        with pytest.raises(Exception):
            halo.exchange()
        halo.close()
    finally:
        for b in blocks:
            b.close()
