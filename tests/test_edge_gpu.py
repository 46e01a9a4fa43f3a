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


def _padded_cavity(mode, dims, pad, dtype):
    """The cavity of cavity_case wrapped in `pad` more void layers, as a block
    of overlap 1 + pad (FreeLB's refined blocks carry overlap 2)."""
    case, p0 = O.cavity_case(mode, *dims, dtype=dtype)
    d = 3 if case.q == 19 else 2
    shp = (case.nz, case.ny, case.nx)
    padw = [(pad, pad) if (a > 0 or d == 3) else (0, 0) for a in range(3)]
    flags = np.pad(case.flags.reshape(shp), padw, constant_values=O.VOID)
    nz, ny, nx = flags.shape
    N = flags.size
    rho = np.pad(case.rho_field.reshape(shp), padw, constant_values=1.0).reshape(-1)
    u = np.stack([np.pad(c.reshape(shp), padw) for c in case.u_field.reshape(d, -1)]).reshape(d, N)
    lat = O.Lattice(case.q, (nx, ny, nz), flags.reshape(-1), case.tasks, case.bcs, omega=case.omega,
                    rho_field=rho, u_field=u, overlap=1 + pad)
    w = O.weights(case.q, dtype)
    pops = np.ascontiguousarray(np.repeat(w[:, None], N, axis=1).astype(dtype))
    return lat, pops


@pytest.mark.parametrize("pad", [1, 2])
@pytest.mark.parametrize("mode,dims", [("hw", (10, 7, 6)), ("hw", (150, 4, 3)), ("fw", (9, 6, 5)),
                                       ("hw", (14, 9, 0))])
def test_overlap_wider_than_one(pad, mode, dims):
    """Blocks whose overlap layer is 2 or 3 cells wide: every cell keeps its
    own slot (the compact row pitch covers the whole row), bit for bit
    against the oracle."""
    case, p0 = _padded_cavity(mode, dims, pad, np.float32)
    print(f"[case] overlap={1 + pad} mode={mode} dims={dims}", flush=True)
    lat = gpu_lattice(case, pops=p0)
    try:
        lat.step(13)
        g = lat.pops()
    finally:
        lat.close()
    ref = case.run(p0.copy(), 13)
    sel = case.flags == O.AABB if mode == "fw" else (case.flags != O.VOID) & interior_mask(case, 1 + pad)
    # fullway: the fluid is the inside of the wall shell; halfway: every cavity cell
    inner = 2 if mode == "fw" else 0
    assert sel.sum() == np.prod([v - inner for v in dims if v > 0])
    assert np.array_equal(g[:, sel], ref[:, sel]), f"max |diff| {np.abs(g[:, sel] - ref[:, sel]).max()}"


def test_watchdog_stops_on_a_nan_population():
    """The NaN / Inf watchdog: a NaN planted in one fluid population from the
    host makes flb_step fail with FLB_ERR_NONFINITE at the first check; a
    healthy run passes every check."""
    from freelb_amd import FLBError, abi
    case, p0 = O.cavity_case("hw", 10, 9, 8)
    lat = gpu_lattice(case, pops=p0)
    try:
        lat.set_watchdog(3)
        lat.step(12)
        p = lat.pops()
        cell = int(np.flatnonzero(case.flags == O.AABB)[17])
        p[5, cell] = np.nan
        lat.upload_pops(p)
        with pytest.raises(FLBError) as e:
            lat.step(6)
        assert e.value.code == abi.FLB_ERR_NONFINITE
        assert "non-finite" in str(e.value)
        assert lat.info().step == 15
        lat.set_watchdog(0)
        lat.step(2)  # switched off: no check
    finally:
        lat.close()
