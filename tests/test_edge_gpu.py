"""Degenerate blocks against the oracle: the smallest cavities (a single
cell, one-cell-thin slabs, no fluid cell at all, only walls), 2-D strips one
cell wide, rows of odd length, and zero-step calls. Bar as in the parity
tests: bit for bit on every cell the reference determines."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_case import gpu_lattice, interior_mask

pytestmark = pytest.mark.gpu


def _both(mode, dims, steps, storage="f32"):
    dtype = np.float64 if storage == "f64" else np.float32
    case, p0 = O.cavity_case(mode, *dims, dtype=dtype)
    print(f"[case] mode={mode} dims={dims} steps={steps} storage={storage}", flush=True)
    lat = gpu_lattice(case, storage=storage, pops=p0)
    try:
        lat.step(steps)
        g = lat.pops()
    finally:
        lat.close()
    ref = case.run(p0.copy(), steps)
    if mode == "fw":
        sel = case.flags == O.AABB
    else:
        sel = (case.flags != O.VOID) & interior_mask(case)
    return g, ref, sel


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 1, 1), (1, 3, 2), (2, 2, 2), (3, 3, 3), (3, 1, 5), (65, 1, 1)])
@pytest.mark.parametrize("steps", [1, 6, 7])
def test_tiny_3d_halfway(dims, steps):
    g, ref, sel = _both("hw", dims, steps)
    assert sel.any()
    assert np.array_equal(g[:, sel], ref[:, sel])


@pytest.mark.parametrize("dims", [(3, 3, 3), (4, 3, 5), (5, 4, 3)])
def test_tiny_3d_fullway(dims):
    g, ref, sel = _both("fw", dims, 9)
    assert np.array_equal(g[:, sel], ref[:, sel])


@pytest.mark.parametrize("dims", [(1, 1, 0), (1, 9, 0), (9, 1, 0), (63, 2, 0), (129, 3, 0)])
@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_thin_2d_halfway(dims, storage):
    g, ref, sel = _both("hw", dims, 11, storage)
    assert np.array_equal(g[:, sel], ref[:, sel])


@pytest.mark.parametrize("storage", ["f32", "f16s", "f64"])
def test_zero_steps_leave_state(storage):
    case, p0 = O.cavity_case("hw", 5, 4, 3, dtype=np.float64 if storage == "f64" else np.float32)
    lat = gpu_lattice(case, storage=storage, pops=p0)
    try:
        before = lat.pops()
        lat.step(0)
        assert lat.info().step == 0
        assert np.array_equal(lat.pops(), before)
    finally:
        lat.close()
