"""The large combined workload at its full size.

1. D3Q19 MRT with Smagorinsky LES on a 512^3 halfway cavity, Re 5000 with
   lid u = 0.05 (tau0 = 0.5 + 3 u N / Re = 0.51536), 5000 steps from rest,
   in FP16-shift and F16C storage against an FP64 run of the same kernel.
   FreeLB has no D3Q19 MRT and no 16-bit storage, so the comparison is with
   FP64 (parity unpinned against the reference). Stated bounds: L2 relative
   velocity error 1e-3 for F16C (measured 6.6e-4) and 4e-3 for IEEE-half
   FP16-shift (measured 2.6e-3); max |d rho| 1e-4 for both.
2. The batched adjoint over 64 seeded channels (64x32x32 .. 128x64x64),
   T = 400 with checkpoints every 20 steps, one launch per step for all
   members: FP32 objective and gradient against the FP64 batch, and the FP32
   gradient of two members against FP64 central differences (h = 1e-4) at
   three seeded cells each, to 1e-3 relative.
"""
import numpy as np
import pytest

import bench

pytestmark = pytest.mark.gpu

TAU0 = 0.5 + 3 * 0.05 * 512 / 5000
BOUNDS = {"f16c": (1e-3, 1e-4), "f16s": (4e-3, 1e-4)}  # (u L2 relative, max |d rho|)


def test_config4_cavity_16bit_vs_fp64():
    print(f"[config4] 512^3 MRT+LES, 5000 steps, tau0 {TAU0}", flush=True)
    res = bench.precision_study(n=512, steps=5000, collisions=("mrt_les",), storages=tuple(BOUNDS),
                                u0_amp=0.0, u_lid=0.05, tau=TAU0)
    print(res, flush=True)
    assert res["runs"]["mrt_les/f64"]["finite"]
    for s, (bu, brho) in BOUNDS.items():
        r = res["runs"][f"mrt_les/{s}"]
        assert r["finite"], s
        assert r["u_l2_rel"] <= bu, (s, r)
        assert r["rho_max_abs"] <= brho, (s, r)


def test_config4_adjoint_batch_vs_fp64_and_fd():
    from freelb_amd.cases import channel_adjoint, channel_batch, channel_dims
    T, K, h, seed = 400, 20, 1e-4, 11
    res = bench.adjoint_batch_bench(n=64, T=T, seed=seed, fp64_check=True)
    print(res, flush=True)
    assert res["grad_finite"]
    assert res["J_max_rel_err_vs_fp64"] <= 1e-5
    assert res["grad_l2_rel_err_vs_fp64"] <= 1e-4
    # FP32 batch gradient of two members against FP64 central differences
    dims = channel_dims(64, seed=seed)
    b, mem, fracs = channel_batch(dims, storage="f32", seed=seed)
    try:
        b.forward(T, K)
        grads = b.backward()
    finally:
        b.close()
        for m in mem:
            m.close()
    rng = np.random.default_rng(5)
    for k in (0, 37):
        g = grads[k]
        cells = rng.choice(g.size, 3, replace=False)
        adj64, s64 = channel_adjoint(dims[k], storage="f64", seed=1000 * seed + k)
        assert np.array_equal(s64, fracs[k])
        with adj64:
            for c in cells:
                adj64.set_fraction_at(int(c), s64[c] + h)
                jp = adj64.forward(T, -1)
                adj64.set_fraction_at(int(c), s64[c] - h)
                jm = adj64.forward(T, -1)
                adj64.set_fraction_at(int(c), s64[c])
                fd = (jp - jm) / (2 * h)
                rel = abs(g[c] - fd) / abs(fd)
                print(f"[config4] member {k} {dims[k]} cell {c}: adjoint {g[c]:.6e} fd {fd:.6e} rel {rel:.2e}",
                      flush=True)
                assert rel <= 1e-3
