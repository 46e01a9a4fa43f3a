"""The oracle's multi-level refined hierarchy (RefinedHierarchy): with one
patch it is the two-level refined cavity that FreeLB's own refined fixtures
pin, bit for bit; with three levels it reproduces FreeLB's own three-level
refined runs (fixtures written by FreeLB, BasicBlock::refine(2)) bit for
bit; the rest state stays a fixed point on every level, and the level
gating steps each block 2^(level) times per coarse step."""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.refine_case import hierarchy_fixture_mismatches, load_fixture, load_fixture3

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "refined3_*.npz"))),
                         ids=lambda p: os.path.basename(p))
def test_three_levels_match_freelb(path):
    fx = load_fixture3(path)
    rh = O.RefinedHierarchy("d2q9", fx["n"], fx["n"], 0, fx["patches"])
    rh.run(fx["counts"])
    bad, checked = hierarchy_fixture_mismatches(rh, fx)
    assert checked > 10000 and bad == 0, (bad, checked)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "refined_*.npz"))),
                         ids=lambda p: os.path.basename(p))
def test_two_level_fixtures_through_the_hierarchy(path):
    fx = load_fixture(path)
    o, m = 1 + fx["p0"] * fx["cell"], (fx["p1"] - fx["p0"]) * fx["cell"]
    rh = O.RefinedHierarchy("d2q9", fx["n"], fx["n"], 0, [(0, (o, o), m)])
    rh.run(fx["counts"])
    bad, checked = hierarchy_fixture_mismatches(rh, fx)
    assert checked > 1000 and bad == 0, (bad, checked)


@pytest.mark.parametrize("lattice,n,cell,p0,p1", [("d2q9", 16, 4, 1, 3), ("d2q9", 24, 4, 2, 4),
                                                  ("d3q19", 12, 4, 1, 2)])
def test_one_patch_is_the_refined_cavity(lattice, n, cell, p0, p1):
    rc = O.RefinedCavity(lattice, n, n, n, cell, p0, p1)
    d = rc.d
    rh = O.RefinedHierarchy(lattice, n, n, n, [(0, rc.origin[:d], rc.m)])
    for k in (1, 4, 9):
        rc.run(k)
        rh.run(k)
        assert np.array_equal(rc.popsC, rh.pops[0]) and np.array_equal(rc.popsF, rh.pops[1])
        assert np.array_equal(rc.coarse.fields(), rh.blocks[0].fields())
        assert np.array_equal(rc.fine.fields(), rh.blocks[1].fields())


@pytest.mark.parametrize("lattice,n,patches", [
    ("d2q9", 40, [(0, (5, 6), 8), (0, (22, 20), (10, 8)), (1, (7, 6), 4)]),
    ("d3q19", 24, [(0, (7, 7, 7), 8), (1, (7, 6, 7), 4)]),
])
def test_rest_state_is_a_fixed_point_on_every_level(lattice, n, patches):
    rh = O.RefinedHierarchy(lattice, n, n, n, patches, u_phys=0.0, dtype=np.float64)
    w = O.weights(rh.q, np.float64)
    rh.run(11)
    for k, p in enumerate(rh.pops):
        live = (rh.blocks[k].flags != O.VOID) & (rh.blocks[k].flags != O.COVERED)
        assert np.abs(p[:, live] - w[:, None]).max() < 1e-14, f"block {k}"


def test_level_gating_counts():
    """Each block of level L steps once every 2^(Lmax - L) counts."""
    rh = O.RefinedHierarchy("d2q9", 40, 40, 0, [(0, (5, 6), 8), (1, (7, 6), 4)])
    steps = [0] * len(rh.blocks)
    for c in range(16):
        for k in range(len(rh.blocks)):
            steps[k] += int(rh.due(rh.levels[k], c))
    assert steps == [4, 8, 16]
