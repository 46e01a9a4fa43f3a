"""Host-side set-up helpers: cavity flags and the vorticity estimator."""
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
