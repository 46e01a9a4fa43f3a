"""Shifted FP16 / BF16 storage against FP64 on the 256^3 cavity.

D3Q19, halfway walls and lid u = 0.1, rho = 1 and a seeded U(-1e-3, 1e-3)
velocity perturbation at the start, 2000 steps, for BGK (tau 0.53) and for
MRT with Smagorinsky LES (C_s 0.17, tau0 0.5005). FreeLB has no reduced-
precision storage, so these results are measured against an FP64 run of the
same kernel (parity unpinned against the reference); the bound on FP16 is
the L2 relative velocity error of 1e-3.
"""
import pytest

import bench

pytestmark = pytest.mark.gpu

BOUNDS = {  # storage: (u L2 relative, max |d rho|, vorticity L2 relative)
    "f32": (1e-5, 1e-6, 1e-4),
    "f16s": (1e-3, 1e-4, 1e-2),
    "bf16s": (3e-2, 3e-3, 2e-1),
}


@pytest.mark.parametrize("collision", ["bgk", "mrt_les"])
def test_shifted_storage_error_vs_fp64(collision):
    res = bench.precision_study(steps=2000, collisions=(collision,))
    print(res)
    ref = res["runs"][f"{collision}/f64"]
    assert ref["finite"]
    for s, (bu, brho, bvort) in BOUNDS.items():
        r = res["runs"][f"{collision}/{s}"]
        assert r["finite"], s
        assert r["u_l2_rel"] <= bu, (s, r)
        assert r["rho_max_abs"] <= brho, (s, r)
        assert r["vort_l2_rel"] <= bvort, (s, r)
