"""D3Q19 MRT and MRT + Smagorinsky on the GPU against the oracle, bit for bit.

These are the collisions of the 256^3 and 512^3 MRT + LES cavities. FreeLB
has no D3Q19 MRT; the oracle's (d'Humieres basis) is pinned by
tests/test_oracle_mrt.py (each moment relaxes at its own rate, the stress
rows take SmagorinskyBGK's omega, a shear wave decays at the BGK viscosity),
and its Smagorinsky omega by FreeLB's own smag fixtures. Here the HIP kernels
must reproduce the oracle bit for bit in FP32 and FP64, in fullway and
halfway cavities, at odd and even step counts, with the default rates and
with a seeded set of distinct rates.

The cavities start from a seeded random velocity field (u ~ U(-1e-3, 1e-3)
per component, rho = 1, f = feq), so every non-hydrodynamic moment is driven
from the first step.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_case import gpu_lattice, interior_mask

pytestmark = pytest.mark.gpu

SHEAR_ROWS = [9, 11, 13, 14, 15]


def _seeded_rates():
    rng = np.random.default_rng(2024)
    r = rng.uniform(0.6, 1.9, 19)
    r[SHEAR_ROWS] = np.nan  # the shear rows take omega (or the LES omega)
    return r


def _start(case, seed=4):
    rng = np.random.default_rng(seed)
    rhou = np.zeros((4, case.N))
    rhou[0] = 1.0
    rhou[1:] = rng.uniform(-1e-3, 1e-3, (3, case.N))
    dtype = np.float64 if case.dtype == "f64" else np.float32
    return case.init_feq(rhou.astype(dtype))


def _run(mode, dims, steps, storage, rates, RT):
    dtype = np.float64 if storage == "f64" else np.float32
    case, _ = O.cavity_case(mode, *dims, RT=RT, dtype=dtype, rates=rates)
    case.dtype = storage
    p0 = _start(case)
    print(f"[mrt case] mode={mode} dims={dims} steps={steps} storage={storage} "
          f"rates={'seeded' if rates is not None else 'default'} RT={RT}", flush=True)
    lat = gpu_lattice(case, storage=storage, pops=p0)
    lat.step(steps)
    g = lat.pops()
    gm = lat.moments(0xFF & ~O.VOID)
    lat.close()
    ref = case.run(p0.copy(), steps)
    rm = case.rhou(ref)
    if mode.endswith("_hw"):
        sel = (case.flags != O.VOID) & interior_mask(case)
    else:
        sel = case.flags == O.AABB
    return g, ref, gm, rm, sel


@pytest.mark.parametrize("storage", ["f32", "f64"])
@pytest.mark.parametrize("mode", ["mrt", "mrt_hw", "smag_mrt", "smag_mrt_hw"])
@pytest.mark.parametrize("steps", [23, 24])
@pytest.mark.parametrize("rates", ["default", "seeded"])
def test_d3q19_mrt_bitwise(storage, mode, steps, rates):
    rt = 0.5005 if mode.startswith("smag") else 0.53
    g, ref, gm, rm, sel = _run(mode, (12, 10, 9), steps, storage,
                               None if rates == "default" else _seeded_rates(), rt)
    assert np.array_equal(g[:, sel], ref[:, sel]), f"max |diff| {np.abs(g[:, sel] - ref[:, sel]).max()}"
    assert np.array_equal(gm[:, sel], rm[:, sel])


@pytest.mark.parametrize("mode", ["mrt_hw", "smag_mrt"])
@pytest.mark.parametrize("dims", [(70, 5, 4), (130, 3, 3)])
def test_d3q19_mrt_bitwise_long_rows(mode, dims):
    """Rows longer than one wavefront, and of lengths that leave lanes idle."""
    g, ref, gm, rm, sel = _run(mode, dims, 17, "f32", None, 0.5005 if mode.startswith("smag") else 0.53)
    assert np.array_equal(g[:, sel], ref[:, sel]), f"max |diff| {np.abs(g[:, sel] - ref[:, sel]).max()}"


def test_d3q19_mrt_les_is_not_bgk():
    """The MRT + LES run departs from plain BGK on the same cavity (the test
    above would also pass if both sides silently ran BGK)."""
    g1, _, _, _, sel = _run("smag_mrt_hw", (12, 10, 9), 24, "f32", None, 0.5005)
    case, _ = O.cavity_case("hw", 12, 10, 9, RT=0.5005)
    case.dtype = "f32"
    p0 = _start(case)
    lat = gpu_lattice(case, pops=p0)
    lat.step(24)
    g2 = lat.pops()
    lat.close()
    assert np.abs(g1[:, sel] - g2[:, sel]).max() > 1e-5


@pytest.mark.parametrize("storage,quantum", [("f16s", 2e-5), ("f16c", 1e-5)])
@pytest.mark.parametrize("mode", ["mrt_hw", "smag_mrt_hw"])
@pytest.mark.parametrize("rates", ["default", "seeded"])
def test_d3q19_mrt_16bit_moment_form_tracks_fp64(storage, quantum, mode, rates):
    """The 16-bit kinds run MRT (+ LES) on the moments of the deviations f - w,
    with the equilibrium moments in closed form and the Smagorinsky stress
    read from the stress rows: arithmetic of their own, not the oracle's.
    From one 16-bit start (rho 1 +- 0.01, u up to 0.08 per component, a random
    non-equilibrium part of 1e-3 w), the GPU must stay within the 16-bit
    rounding of each stored step (`quantum` per step) of the FP64 oracle over
    one and three steps. A wrong closed-form term or stress row moves the
    populations by ~1e-3 at these speeds."""
    rt = 0.5005 if mode.startswith("smag") else 0.53
    rt_rates = None if rates == "default" else _seeded_rates()
    case, _ = O.cavity_case(mode, 12, 10, 9, RT=rt, dtype=np.float64, rates=rt_rates)
    rng = np.random.default_rng(17)
    rhou = np.zeros((4, case.N))
    rhou[0] = 1.0 + rng.uniform(-0.01, 0.01, case.N)
    rhou[1:] = rng.uniform(-0.08, 0.08, (3, case.N))
    p = case.init_feq(rhou)
    p = p * (1.0 + rng.uniform(-1e-3, 1e-3, p.shape))
    print(f"[mrt 16-bit] storage={storage} mode={mode} rates={rates}", flush=True)
    lat = gpu_lattice(case, storage=storage, pops=p.astype(np.float32))
    try:
        ref = lat.pops().astype(np.float64)  # the quantised start
        sel = (case.flags != O.VOID) & interior_mask(case)
        done = 0
        for steps in (1, 3):
            lat.step(steps - done)
            ref = case.run(ref, steps - done)
            done = steps
            err = float(np.abs(lat.pops()[:, sel].astype(np.float64) - ref[:, sel]).max())
            print(f"  after {steps} steps: max |GPU - FP64| = {err:.3e}", flush=True)
            assert err <= quantum * steps, (steps, err)
    finally:
        lat.close()
