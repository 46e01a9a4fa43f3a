"""Shifted FP16 / BF16 / F16C storage against FP64 on the 256^3 cavity.

D3Q19, halfway walls and lid u = 0.1, rho = 1 and a seeded U(-1e-3, 1e-3)
velocity perturbation at the start, 2000 steps, for BGK (tau 0.53) and for
MRT with Smagorinsky LES (C_s 0.17, tau0 0.5005). FreeLB has no reduced-
precision storage, so these results are measured against an FP64 run of the
same kernel (parity unpinned against the reference).

The target for 16-bit storage is an L2 relative velocity error of at most
1e-3 on this BGK case: both IEEE-half shift (f16s) and F16C meet it, and the
test holds them to it. The other bounds are regression guards about twice
the measured errors.
"""
import pytest

import bench

pytestmark = pytest.mark.gpu

# storage: (u L2 relative, max |d rho|, vorticity L2 relative), per collision;
# measured on MI355X (the shifted kinds' BGK runs on the deviations f - w):
# BGK f32 7.3e-5 / 5.8e-5 / 7.7e-5, f16s 2.9e-4 / 8.9e-5 / 1.1e-3,
# bf16s 5.5e-3 / 6.0e-4 / 9.7e-3, f16c 1.4e-4 / 3.8e-5 / 5.5e-4;
# MRT+LES f32 3.4e-4, f16s 3.7e-3, bf16s 5.6e-2, f16c 1.3e-3 (u).
TARGET_U = 1e-3
BOUNDS = {
    "bgk": {"f32": (2e-4, 2e-4, 2e-4), "f16s": (TARGET_U, 3e-4, 2.5e-3), "bf16s": (1.2e-2, 1.5e-3, 2e-2),
            "f16c": (TARGET_U, 1e-4, 1.2e-3)},
    "mrt_les": {"f32": (1e-3, 1e-4, 1e-3), "f16s": (8e-3, 2e-3, 8e-3), "bf16s": (1.2e-1, 1e-2, 8e-2),
                "f16c": (3e-3, 1e-3, 3e-3)},
}


@pytest.mark.parametrize("collision", ["bgk", "mrt_les"])
def test_shifted_storage_error_vs_fp64(collision):
    res = bench.precision_study(steps=2000, collisions=(collision,))
    print(res)
    ref = res["runs"][f"{collision}/f64"]
    assert ref["finite"]
    for s, (bu, brho, bvort) in BOUNDS[collision].items():
        r = res["runs"][f"{collision}/{s}"]
        assert r["finite"], s
        assert r["u_l2_rel"] <= bu, (s, r)
        assert r["rho_max_abs"] <= brho, (s, r)
        assert r["vort_l2_rel"] <= bvort, (s, r)
