"""Host-side set-up helpers: cavity flags, the vorticity estimator and the
CPU baseline of bench.py."""
import os

import numpy as np
import pytest

import bench
from freelb_amd import cases
from oracle import oracle as O


@pytest.mark.parametrize("dims", [(8, 8, 8), (8, 7, 6), (100, 100, 100)])
def test_cavity_flags_match_oracle(dims):
    assert np.array_equal(cases.cavity_flags(*dims), O.cavity_flags(*dims))


def test_vorticity_of_linear_fields():
    n = 12
    m = n + 2
    z, y, x = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    rhou = np.zeros((4, m ** 3))
    # u = (-0.3 y + 0.1 z, 0.5 x, 0.2 y) has curl (0.2, 0.1, 0.8)
    rhou[1] = (-0.3 * y + 0.1 * z).ravel()
    rhou[2] = (0.5 * x).ravel()
    rhou[3] = (0.2 * y).ravel()
    w = bench._vorticity(rhou, n)
    assert w.shape == (3, n - 2, n - 2, n - 2)
    np.testing.assert_allclose(w[0], 0.2, atol=1e-12)
    np.testing.assert_allclose(w[1], 0.1, atol=1e-12)
    np.testing.assert_allclose(w[2], 0.8, atol=1e-12)


def test_cpu_baseline_times_the_reference_harness():
    """bench.py's cpu_baseline runs FreeLB's halfway cavity built from the
    reference sources (oracle/_ref/cavity_hw_bench) and reads its result line
    past FreeLB's own geometry printout; the C oracle is only the fallback."""
    exe = os.path.join(os.path.dirname(bench.__file__), "oracle", "_ref", "cavity_hw_bench")
    if not os.path.exists(exe):
        pytest.skip("reference harness not built here")
    rec = bench.cpu_baseline(n=24, steps=4, warmup=1)
    assert rec["kind"] == "reference", rec
    assert rec["value"] > 0 and rec["unit"] == "MLUPS"


def test_adjoint_byte_accounting():
    """Bytes behind the adjoint GB/s figures. T = 800, K = 28: 29 segments,
    so 771 recompute steps of 172 B and 29 moment records of 96 B beside the
    800 adjoint steps of 180 B; K = 0 recomputes nothing and its forward steps
    write the 16-byte moment record too."""
    cells, T = 1000, 800
    r = bench._adjoint_rates(cells, T, 28, 1.0, 1.0)
    assert r["recomputed_steps"] == 771
    assert r["backward_gbs"] == round(cells * (800 * 180 + 771 * 172 + 29 * 96) / 1e6, 1)
    assert r["forward_gbs"] == round(cells * 800 * 156 / 1e6, 1)
    r = bench._adjoint_rates(cells, T, 0, 1.0, 1.0)
    assert r["recomputed_steps"] == 0
    assert r["backward_gbs"] == round(cells * 800 * 180 / 1e6, 1)
    assert r["forward_gbs"] == round(cells * 800 * 172 / 1e6, 1)
