"""Batched stepping (one launch per step over a block table) against the same
blocks stepped one by one: the batch runs the very same per-cell update, so
populations must agree bit for bit, through the persistent grid, the hipGraph
replay, the autotuner and settings changed after the batch was built."""
import numpy as np
import pytest

from freelb_amd import BlockLattice, BlockLatticeManager, FLBError, HalfwayBB, Key_TypePair, abi
from freelb_amd.cases import cavity_block

pytestmark = pytest.mark.gpu

SHAPES = [(12, 9, 7), (17, 16, 15), (70, 5, 9), (33, 33, 3), (5, 6, 40), (64, 64, 8)]


def _channel(storage="f32", nx=24, ny=10, nz=7, tau=0.7):
    """Periodic-x channel, halfway walls in y and z (exercises the wrapped kernel)."""
    lat = BlockLattice(19, (nx, ny + 2, nz + 2), omega=1.0 / tau, storage=storage,
                       periodic=(True, False, False))
    f = np.full((nz + 2, ny + 2, nx), 1, dtype=np.uint8)
    f[1:-1, 1:-1, :] = 2
    lat.upload_flags(f.reshape(-1))
    lat.set_tasks([Key_TypePair(2, abi.DYN_BGK)])
    lat.set_halfway_bc([HalfwayBB(2, 1, False)])
    rng = np.random.default_rng(5)
    N = f.size
    u = np.zeros((3, N))
    u[:, f.reshape(-1) == 2] = rng.uniform(-0.02, 0.02, size=(3, int((f == 2).sum())))
    lat.init_equilibrium(np.ones(N), u)
    return lat


def _blocks(storage="f32", mixed=False):
    rng = np.random.default_rng(11)
    out = []
    for k, sh in enumerate(SHAPES):
        coll = "mrt_les" if (mixed and k % 3 == 1) else "bgk"
        out.append(cavity_block(sh, storage=storage, collision=coll, tau=float(rng.uniform(0.51, 0.9)),
                                u_lid=float(rng.uniform(0.02, 0.1)), u0_amp=1e-3, seed=k))
    out.append(_channel(storage))
    return out


def _close(bs):
    for b in bs:
        b.close()


@pytest.mark.parametrize("storage,mixed", [("f32", False), ("f16s", False), ("f16c", True), ("f32", True), ("f64", True)])
def test_batch_step_matches_single_blocks(storage, mixed):
    ref, bat = _blocks(storage, mixed), _blocks(storage, mixed)
    try:
        for b in ref:
            b.step(23)
        with BlockLatticeManager(bat) as m:
            assert m.info()["blocks"] == len(bat)
            m.step(23)
            m.synchronize()
        for r, b in zip(ref, bat):
            assert np.array_equal(r.pops(), b.pops())
    finally:
        _close(ref)
        _close(bat)


@pytest.mark.parametrize("wpb,grid", [(1, 0), (2, 0), (4, 0), (4, 8), (1, 3), (2, 512)])
def test_batch_launch_shapes(wpb, grid):
    ref, bat = _blocks(), _blocks()
    try:
        for b in ref:
            b.step(6)
        with BlockLatticeManager(bat) as m:
            m.set_launch(wpb, grid)
            m.step(6)
            m.synchronize()
        for r, b in zip(ref, bat):
            assert np.array_equal(r.pops(), b.pops())
    finally:
        _close(ref)
        _close(bat)


def test_graph_replay_matches_eager_each_time():
    ref, bat = _blocks(), _blocks()
    try:
        with BlockLatticeManager(bat) as m:
            m.capture(10)
            for rep in range(3):
                for b in ref:
                    b.step(10)
                m.replay()
                m.synchronize()
                # eager work between replays: read-out on the blocks' own streams
                for r, b in zip(ref, bat):
                    assert np.array_equal(r.pops(), b.pops()), f"replay {rep}"
                    assert np.array_equal(r.moments(), b.moments())
            assert bat[0].info().step == 30
    finally:
        _close(ref)
        _close(bat)


def test_autotune_keeps_state_and_picks_a_valid_shape():
    bat = _blocks()
    try:
        with BlockLatticeManager(bat) as m:
            m.step(3)
            m.synchronize()
            before = [b.pops() for b in bat]
            best = m.autotune(4)
            assert best["waves_per_group"] in (1, 2, 4) and best["grid"] >= 0 and best["ms_per_step"] > 0
            assert m.info()["waves_per_group"] == best["waves_per_group"]
            for p, b in zip(before, bat):
                assert np.array_equal(p, b.pops())
    finally:
        _close(bat)


def test_settings_changed_after_creation_are_seen():
    ref, bat = _blocks(), _blocks()
    try:
        with BlockLatticeManager(bat) as m:
            m.step(4)
            for b in ref:
                b.step(4)
            bat[2].set_omega(1.2)
            ref[2].set_omega(1.2)
            bat[1].set_tasks([Key_TypePair(2 | 4 | 8, abi.DYN_MRT)])
            ref[1].set_tasks([Key_TypePair(2 | 4 | 8, abi.DYN_MRT)])
            m.step(6)
            m.synchronize()
            for b in ref:
                b.step(6)
        for r, b in zip(ref, bat):
            assert np.array_equal(r.pops(), b.pops())
    finally:
        _close(ref)
        _close(bat)


def test_batch_rejects_mixed_parity_and_kinds():
    bat = _blocks()
    other = cavity_block(8, storage="f16s")
    try:
        with pytest.raises(FLBError):
            BlockLatticeManager(bat + [other])
        with BlockLatticeManager(bat) as m:
            m.capture(2)  # recorded at parity 0
            m.step(1)
            with pytest.raises(FLBError) as e:
                m.replay()  # blocks now at parity 1
            assert e.value.code == abi.FLB_ERR_STATE
            bat[0].step(1)
            with pytest.raises(FLBError) as e:
                m.step(1)  # block 0 at parity 0, the rest at 1
            assert e.value.code == abi.FLB_ERR_STATE
    finally:
        other.close()
        _close(bat)
