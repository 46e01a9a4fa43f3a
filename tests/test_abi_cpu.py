"""The C ABI without a GPU: the library loads, exports every symbol the header
declares, and its argument checks fail cleanly before any device work."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from freelb_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "flb_abi.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(flb_\w+)\s*\(", text)))


def test_header_symbols_exported():
    lib = C.CDLL(abi.LIB_PATH)
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared in flb_abi.h but not exported: {missing}"
    assert set(names) == set(abi.EXPORTS), "abi.EXPORTS out of step with the header"


def test_abi_version():
    assert abi.lib().flb_abi_version() == 1


def _desc(**kw):
    d = abi.BlockDesc()
    d.q, d.storage, d.nx, d.ny, d.nz, d.overlap, d.omega = 19, 0, 10, 10, 10, 1, 1.5
    for k, v in kw.items():
        setattr(d, k, v)
    return d


@pytest.mark.parametrize("field,value,msg", [
    ("q", 15, "q must be"),
    ("storage", 9, "storage kind"),
    ("nx", 0, "extents"),
    ("omega", 2.5, "omega"),
    ("overlap", 0, "overlap"),
])
def test_create_rejects_bad_descriptor(field, value, msg):
    L = abi.lib()
    h = C.c_void_p()
    rc = L.flb_block_create(C.byref(_desc(**{field: value})), C.byref(h))
    assert rc == abi.FLB_ERR_ARG
    assert msg in L.flb_last_error().decode()
    assert not h.value


def test_create_rejects_d2q9_with_depth():
    L = abi.lib()
    h = C.c_void_p()
    assert L.flb_block_create(C.byref(_desc(q=9, nz=4)), C.byref(h)) == abi.FLB_ERR_ARG


def test_null_arguments():
    L = abi.lib()
    assert L.flb_block_create(None, None) == abi.FLB_ERR_ARG
    for fn in ("flb_block_destroy",):
        assert getattr(L, fn)(None) == abi.FLB_OK  # destroying nothing is a no-op
    for fn in ("flb_synchronize", "flb_apply_halfway_bc"):
        assert getattr(L, fn)(None) == abi.FLB_ERR_ARG
    for fn in ("flb_apply_cell_dynamics", "flb_stream"):
        assert getattr(L, fn)(None, 0) == abi.FLB_ERR_ARG
    assert L.flb_set_level(None, 0, 0) == abi.FLB_ERR_ARG
    assert L.flb_set_watchdog(None, 1) == abi.FLB_ERR_ARG
    for fn in ("flb_refine_average", "flb_refine_interpolate", "flb_refine_communicate"):
        assert getattr(L, fn)(None, 0) == abi.FLB_ERR_ARG
    assert L.flb_refine_set_levels(None, 0, 1) == abi.FLB_ERR_ARG
    assert L.flb_reduce_mass_momentum_noalloc(None, 0xFE, None) == abi.FLB_ERR_ARG
    assert L.flb_step(None, 1) == abi.FLB_ERR_ARG
    assert L.flb_set_tasks(None, None, 0) == abi.FLB_ERR_ARG
    assert "null block" in L.flb_last_error().decode()
    g = C.c_double()
    assert L.flb_bandwidth_probe(0, 1, C.byref(g)) == abi.FLB_ERR_ARG


def test_no_gpu_fails_loudly():
    """Without a device the create call reports a HIP error (there is no CPU path)."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    from freelb_amd import BlockLattice, FLBError
    with pytest.raises(FLBError) as e:
        BlockLattice(19, (8, 8, 8), omega=1.5)
    assert e.value.code == abi.FLB_ERR_HIP


def test_host_mirror_shapes():
    from freelb_amd import HalfwayBB, Key_TypePair, weights
    assert Key_TypePair(2, abi.DYN_BGK).smagorinsky == 0.0
    assert HalfwayBB(4).void_mask == 1 and not HalfwayBB(4).moving
    w = weights(19)
    assert abs(w.sum() - 1.0) < 1e-15 and w.shape == (19,)
    assert abs(weights(9).sum() - 1.0) < 1e-15


def test_batch_null_arguments():
    L = abi.lib()
    h = C.c_void_p()
    assert L.flb_batch_create(None, 0, None, C.byref(h)) == abi.FLB_ERR_ARG
    assert L.flb_batch_destroy(None) == abi.FLB_OK
    for fn in ("flb_batch_replay", "flb_batch_synchronize"):
        assert getattr(L, fn)(None) == abi.FLB_ERR_ARG
    assert L.flb_batch_step(None, 1) == abi.FLB_ERR_ARG
    assert L.flb_batch_capture(None, 2) == abi.FLB_ERR_ARG
    assert L.flb_batch_set_launch(None, None) == abi.FLB_ERR_ARG
    assert L.flb_batch_autotune(None, 2, None, None) == abi.FLB_ERR_ARG
    assert "null batch" in L.flb_last_error().decode()


def test_adjoint_batch_null_arguments():
    L = abi.lib()
    h = C.c_void_p()
    assert L.flb_adjoint_batch_create(None, 1, None, C.byref(h)) == abi.FLB_ERR_ARG
    arr = (C.c_void_p * 2)(None, None)
    assert L.flb_adjoint_batch_create(arr, 0, None, C.byref(h)) == abi.FLB_ERR_ARG
    assert "at least one member" in L.flb_last_error().decode()
    assert L.flb_adjoint_batch_create(arr, 2, None, C.byref(h)) == abi.FLB_ERR_ARG
    assert "member 0 is null" in L.flb_last_error().decode()
    assert not h.value
    assert L.flb_adjoint_batch_destroy(None) == abi.FLB_OK
    J = C.c_double()
    assert L.flb_adjoint_batch_forward(None, 1, -1, C.byref(J)) == abi.FLB_ERR_ARG
    assert L.flb_adjoint_batch_backward(None, C.byref(J)) == abi.FLB_ERR_ARG
    assert L.flb_adjoint_batch_timing(None, None, None) == abi.FLB_ERR_ARG
    assert "null batch" in L.flb_last_error().decode()


def test_output_null_arguments():
    L = abi.lib()
    assert L.flb_write_vtm(None, 1, None, None, None, b".", b"x", 0, None, None, None) == abi.FLB_ERR_ARG
    assert L.flb_store_moments(None, 0xFE) == abi.FLB_ERR_ARG
    r = C.c_double()
    assert L.flb_tolerance(None, 1, 0, C.byref(r)) == abi.FLB_ERR_ARG
    assert L.flb_enable_tolerance(None) == abi.FLB_ERR_ARG


def _layout(**kw):
    lay = abi.Layout()
    abi.check(abi.lib().flb_block_layout(C.byref(_desc(**kw)), C.byref(lay)))
    return lay


@pytest.mark.parametrize("overlap", [1, 2, 3])
@pytest.mark.parametrize("nx", [7, 12, 40, 96, 200, 224, 225, 226, 230, 300, 520])
@pytest.mark.parametrize("periodic_x", [0, 1])
def test_slots_are_distinct_for_every_overlap(overlap, nx, periodic_x):
    """Every host cell gets its own flag / field slot and its own population
    slot per direction, for compact (short) and wide rows and overlap
    widths 1-3 (FreeLB's refined blocks use 2); slots stay inside the
    allocation with the front pad before them."""
    per = (C.c_int * 3)(periodic_x, 0, 0)
    ny, nz = 2 * overlap + 3, 2 * overlap + 2
    lay = _layout(nx=nx, ny=ny, nz=nz, overlap=overlap, periodic=per)
    x = np.arange(nx)[None, None, :]
    y = np.arange(ny)[None, :, None]
    z = np.arange(nz)[:, None, None]
    assert lay.xs == (0 if periodic_x else overlap)
    slots = ((x - lay.xs) + lay.pitch * y + lay.plane * z).reshape(-1)
    assert np.unique(slots).size == slots.size
    assert slots.min() + lay.front_pad >= 0 and slots.max() + lay.front_pad < lay.stride
    pops = ((x - lay.xs) + lay.ypop * y + lay.zpop * z).reshape(1, -1) + lay.dpop * np.arange(19)[:, None]
    assert np.unique(pops).size == pops.size
    assert pops.min() + lay.front_pad >= 0 and pops.max() + lay.front_pad < lay.pop_elems
    # the odd step reads and writes one cell past an updated cell: still inside
    assert pops.max() + 1 + lay.front_pad < lay.pop_elems


def test_layout_rejects_like_create():
    L = abi.lib()
    lay = abi.Layout()
    assert L.flb_block_layout(C.byref(_desc(overlap=0)), C.byref(lay)) == abi.FLB_ERR_ARG
    assert "overlap" in L.flb_last_error().decode()


@pytest.mark.parametrize("dims,storage", [
    ((2048, 2048, 2048), "f32"),   # 2^33 cells
    ((1025, 1024, 1024), "f32"),   # just over 2^30 cells: 4 B x N > 2^32
    ((513, 1024, 1024), "f64"),    # just over 2^29 cells: 8 B x N > 2^32
])
def test_adjoint_rejects_boxes_past_32bit_offsets(dims, storage):
    """The adjoint steps address a population array with 32-bit byte offsets:
    creation refuses a box whose arrays exceed 2^32 bytes, before any device
    work."""
    L = abi.lib()
    d = abi.AdjointDesc()
    d.nx, d.ny, d.nz = dims
    d.storage = abi.STORAGE[storage]
    d.omega = 1.0
    d.outlet_x = 0
    h = C.c_void_p()
    assert L.flb_adjoint_create(C.byref(d), C.byref(h)) == abi.FLB_ERR_ARG
    assert "box too large" in L.flb_last_error().decode()
    assert not h.value
