"""The pair form of the step kernel (16-bit storage, two cells per lane,
word-wide accesses with lane shifts at the odd step) against the cell form:
the same per-cell arithmetic, so the populations must agree bit for bit in
every mode, at odd row lengths, with periodic y/z axes and with dead cells."""
import numpy as np
import pytest

from freelb_amd import BlockLattice, HalfwayBB, Key_TypePair, abi
from freelb_amd.cases import cavity_block
from gpu_case import gpu_lattice
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _pair(monkeypatch, make):
    monkeypatch.setenv("FLB_LAUNCH", "-2,4")
    try:
        return make()
    finally:
        monkeypatch.delenv("FLB_LAUNCH")


@pytest.mark.parametrize("storage", ["f16s", "bf16s"])
@pytest.mark.parametrize("mode", ["fw", "hw", "hwfield", "mrt", "smag"])
@pytest.mark.parametrize("dims", [(8, 7, 6), (131, 5, 4), (257, 3, 3), (64, 9, 5)])
def test_pair_form_matches_cell_form(monkeypatch, storage, mode, dims):
    case, p = O.cavity_case(mode, *dims, dtype=np.float32)
    rng = np.random.default_rng(sum(dims))
    p = (p * (1 + 0.01 * rng.standard_normal(p.shape))).astype(np.float32)
    cell = gpu_lattice(case, storage, p)
    pair = _pair(monkeypatch, lambda: gpu_lattice(case, storage, p))
    try:
        for n in (1, 2, 7):
            cell.step(n)
            pair.step(n)
            assert np.array_equal(cell.pops(), pair.pops()), (mode, dims, n)
    finally:
        cell.close()
        pair.close()


def test_pair_form_big_cavity(monkeypatch):
    a = cavity_block((300, 40, 30), storage="f16s", u0_amp=1e-3, seed=3)
    b = _pair(monkeypatch, lambda: cavity_block((300, 40, 30), storage="f16s", u0_amp=1e-3, seed=3))
    try:
        a.step(41)
        b.step(41)
        assert np.array_equal(a.pops(), b.pops())
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("storage", ["f16s", "bf16s"])
def test_pair_form_periodic_yz_channel(monkeypatch, storage):
    nx, ny, nz = 70, 12, 9

    def make():
        lat = BlockLattice(19, (nx + 2, ny, nz), omega=1.0 / 0.7, storage=storage,
                           periodic=(False, True, True))
        f = np.full((nz, ny, nx + 2), 1, dtype=np.uint8)
        f[:, :, 1:-1] = 2
        f[3:5, 4:7, 20:23] = 4  # a solid block inside (not updated)
        lat.upload_flags(f.reshape(-1))
        lat.set_tasks([Key_TypePair(2, abi.DYN_BGK)])
        lat.set_halfway_bc([HalfwayBB(2, 1 | 4, False)])
        N = f.size
        rng = np.random.default_rng(1)
        u = np.zeros((3, N))
        u[:, f.reshape(-1) == 2] = rng.uniform(-0.03, 0.03, size=(3, int((f == 2).sum())))
        lat.init_equilibrium(np.ones(N), u)
        return lat

    a = make()
    b = _pair(monkeypatch, make)
    try:
        a.step(13)
        b.step(13)
        assert np.array_equal(a.pops(), b.pops())
    finally:
        a.close()
        b.close()
