"""Two-level refinement on the GPU (flb_refine_* through the C ABI): the
refined cavity against the oracle bit for bit in FP32 and FP64 (D2Q9 at the
sizes of FreeLB's own refined-cavity fixtures, which the oracle reproduces,
and D3Q19), at odd and even counts, populations and RHO / VELOCITY fields;
the 16-bit storage kinds against FP64; and the argument checks."""
import numpy as np
import pytest

from freelb_amd import BlockLattice, Key_TypePair, Refinement
from oracle import oracle as O
from refine_case import (compare_masks, fixture_mismatches, gpu_hierarchy, gpu_refined, hierarchy_fixture_mismatches,
                         hierarchy_masks, load_fixture, load_fixture3)

pytestmark = pytest.mark.gpu

CASES = [("d2q9", 16, 4, 1, 3), ("d2q9", 24, 4, 2, 4), ("d3q19", 12, 4, 1, 2), ("d3q19", 16, 4, 1, 3)]


def _check(rc, coarse, fine, exact=True, tol=0.0):
    cm, fm = compare_masks(rc)
    pc, pf = coarse.pops(), fine.pops()
    if exact:
        assert np.array_equal(pc[:, cm], rc.popsC[:, cm]), "coarse populations differ"
        assert np.array_equal(pf[:, fm], rc.popsF[:, fm]), "fine populations differ"
        live = rc.coarse.flags == O.AABB
        assert np.array_equal(coarse.download_fields()[:, live], rc.coarse.fields()[:, live])
        fin = np.zeros(rc.fdims[::-1], dtype=bool)
        if rc.d == 3:
            fin[2:-2, 2:-2, 2:-2] = True
        else:
            fin[0, 2:-2, 2:-2] = True
        fin = fin.reshape(-1)
        assert np.array_equal(fine.download_fields()[:, fin], rc.fine.fields()[:, fin])
    else:
        assert np.abs(pc[:, cm] - rc.popsC[:, cm]).max() <= tol
        assert np.abs(pf[:, fm] - rc.popsF[:, fm]).max() <= tol


@pytest.mark.parametrize("lattice,n,cell,p0,p1", CASES)
@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_refined_cavity_matches_oracle(lattice, n, cell, p0, p1, storage):
    dt = np.float32 if storage == "f32" else np.float64
    rc = O.RefinedCavity(lattice, n, n, n, cell, p0, p1, dtype=dt)
    coarse, fine, ref = gpu_refined(rc, storage)
    try:
        info = ref.info()
        assert info["average"] > 0 and info["interp"] > 0
        for k in (1, 1, 3, 6, 11):
            rc.run(k)
            ref.run(k)
            _check(rc, coarse, fine)
    finally:
        ref.close()
        coarse.close()
        fine.close()


# Hierarchies of several patches and levels: (lattice, n, patches), patches as
# (parent block, origin in the parent's cells, extent in parent cells)
HIERARCHIES = [
    # two level-1 patches side by side, a level-2 patch inside the first
    ("d2q9", 40, [(0, (5, 6), 8), (0, (22, 20), (10, 8)), (1, (7, 6), 4)]),
    # three levels nested in 3D
    ("d3q19", 24, [(0, (7, 7, 7), 8), (1, (7, 6, 7), 4)]),
    # two 3D level-1 patches, one with a level-2 child
    ("d3q19", 28, [(0, (4, 5, 6), 6), (0, (17, 15, 14), 6), (2, (5, 5, 5), 2)]),
]


@pytest.mark.parametrize("lattice,n,patches", HIERARCHIES, ids=["2d-two-patches-3-levels", "3d-3-levels",
                                                                "3d-two-patches-3-levels"])
@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_refined_hierarchy_matches_oracle(lattice, n, patches, storage):
    """Several patches and three levels (each level steps every 2^(Lmax - L)
    counts, averages before interpolations): every block bit for bit
    against the oracle's hierarchy at odd and even counts."""
    dt = np.float32 if storage == "f32" else np.float64
    rh = O.RefinedHierarchy(lattice, n, n, n, patches, dtype=dt)
    print(f"[hierarchy] {lattice} {n} {patches} {storage}", flush=True)
    blocks, refs, mgr = gpu_hierarchy(rh, storage)
    masks = hierarchy_masks(rh)
    try:
        for k in (1, 2, 3, 7, 13):
            rh.run(k)
            mgr.run(k)
            for j, (b, m) in enumerate(zip(blocks, masks)):
                assert np.array_equal(b.pops()[:, m], rh.pops[j][:, m]), f"block {j} after {rh.count} counts"
                fl = rh.blocks[j].flags
                live = m & (fl == O.AABB)
                assert np.array_equal(b.download_fields()[:, live], rh.blocks[j].fields()[:, live]), f"fields {j}"
    finally:
        for r in refs:
            r.close()
        for b in blocks:
            b.close()


@pytest.mark.parametrize("name", ["refined3_d2q9_24_cell4_p15_q24_37.npz", "refined3_d2q9_32_cell4_p17_q35_44.npz"])
def test_three_level_hierarchy_matches_freelb_fixture(name):
    """Three levels on the GPU against FreeLB's own three-level run: the GPU
    equals the oracle hierarchy bit for bit, and the oracle equals FreeLB's
    blocks (the oracle is the map between the two block layouts)."""
    import os
    fx = load_fixture3(os.path.join(os.path.dirname(__file__), "golden", name))
    rh = O.RefinedHierarchy("d2q9", fx["n"], fx["n"], 0, fx["patches"])
    blocks, refs, mgr = gpu_hierarchy(rh)
    masks = hierarchy_masks(rh)
    try:
        rh.run(fx["counts"])
        mgr.run(fx["counts"])
        for j, (b, m) in enumerate(zip(blocks, masks)):
            assert np.array_equal(b.pops()[:, m], rh.pops[j][:, m]), f"block {j}"
        bad, checked = hierarchy_fixture_mismatches(rh, fx)
        assert checked > 10000 and bad == 0
    finally:
        for r in refs:
            r.close()
        for b in blocks:
            b.close()


def test_refined_cavity_matches_freelb_fixture():
    """The GPU against FreeLB's own refined run (the oracle is only the map
    between the two block layouts: it must match the fixture too)."""
    import glob
    import os
    path = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "refined_*.npz")))[0]
    fx = load_fixture(path)
    rc = O.RefinedCavity("d2q9", fx["n"], fx["n"], 1, fx["cell"], fx["p0"], fx["p1"])
    coarse, fine, ref = gpu_refined(rc, "f32")
    try:
        ref.run(fx["counts"])
        cm, fm = compare_masks(rc)
        rc.popsC[:, cm] = coarse.pops()[:, cm]
        rc.popsF[:, fm] = fine.pops()[:, fm]
        rc.coarse.buffers["rho_field"][:] = coarse.download_fields()[0]
        rc.coarse.buffers["u_field"][:] = coarse.download_fields()[1:].reshape(-1)
        rc.fine.buffers["rho_field"][:] = fine.download_fields()[0]
        rc.fine.buffers["u_field"][:] = fine.download_fields()[1:].reshape(-1)
        # the fine block's outer ghost layer is never read by an updated cell
        # and not updated here, so it is left out
        bad, checked = fixture_mismatches(rc, fx, fine_mask=fm)
        assert checked > 0 and bad == 0, f"{bad} of {checked} cells differ"
    finally:
        ref.close()
        coarse.close()
        fine.close()


@pytest.mark.parametrize("storage,tol", [("f16s", 2e-4), ("bf16s", 2e-3), ("f16c", 2e-4)])
def test_refined_cavity_16bit_close_to_fp64(storage, tol):
    rc64 = O.RefinedCavity("d3q19", 12, 12, 12, 4, 1, 2, dtype=np.float64)
    rc = O.RefinedCavity("d3q19", 12, 12, 12, 4, 1, 2, dtype=np.float64)
    coarse, fine, ref = gpu_refined(rc, storage)
    try:
        rc64.run(20)
        ref.run(20)
        _check(rc64, coarse, fine, exact=False, tol=tol)
    finally:
        ref.close()
        coarse.close()
        fine.close()


def test_refine_rejects_bad_layouts():
    rc = O.RefinedCavity("d3q19", 12, 12, 12, 4, 1, 2)
    coarse, fine, ref = gpu_refined(rc, "f32")
    ref.close()
    try:
        with pytest.raises(Exception):
            Refinement(coarse, coarse, rc.origin)
        with pytest.raises(Exception):  # stencil would leave the coarse block
            Refinement(coarse, fine, (1, 1, 1))
        small = BlockLattice(19, (11, 12, 12), omega=float(rc.omega_f))  # 11 is not 2m + 4
        fl = np.full((12, 12, 11), O.COVERED, dtype=np.uint8)
        fl[1:-1, 1:-1, 1:-1] = O.AABB
        small.upload_flags(fl.reshape(-1))
        small.set_tasks([Key_TypePair(O.AABB, O.DYN_BGK_WRITE)])
        with pytest.raises(Exception):
            Refinement(coarse, small, rc.origin)
        small.close()
    finally:
        coarse.close()
        fine.close()
