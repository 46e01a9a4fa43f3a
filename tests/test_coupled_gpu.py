"""Coupled flow + temperature lattices on the GPU (flb_share_velocity,
flb_add_buoyancy, FLB_DYN_BGK_SCALAR_FORCE, the D2Q5 / D3Q7 blocks and the
anti-bounce-back walls, through the C ABI).

  * FreeLB's own coupled run (tests/golden/natconv_*.npz): FP32 populations
    of both lattices, RHO / VELOCITY / TEMP bit for bit;
  * the oracle at odd and even step counts in FP32 and FP64, bit for bit,
    also on a 48^3 cavity and with two scalar lattices on one flow block;
  * at a large size (1024^2 D2Q9 + D2Q5 in FP64, closed adiabatic box), the
    scalar sum is conserved to rounding;
  * argument checks and the detaching of couplings when a block is destroyed.
"""
import os

import numpy as np
import pytest

from freelb_amd import BlockLatManagerCoupling, BlockLattice, FLBError, HalfwayBB, Key_TypePair, abi
from natconv_case import ALL, AABB, BB, COLD, HOT, VOID, fixtures, gpu_pair, load, oracle_run
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("path", fixtures(), ids=os.path.basename)
def test_coupled_matches_freelb_fixture(path):
    fx = load(path)
    flow, temp, cpl = gpu_pair(fx, "f32")
    try:
        n = int(fx["steps"])
        cpl.step(1)
        cpl.step(n - 1)
        # every updated cell (the void shell is the blocks' overlap layer, which
        # the GPU neither updates nor reads back)
        live = fx["flags"] != VOID
        assert np.array_equal(flow.pops()[:, live].view(np.uint32), fx["flow"][:, live].view(np.uint32))
        assert np.array_equal(temp.pops()[:, live].view(np.uint32), fx["temp"][:, live].view(np.uint32))
        g = flow.download_fields()
        assert np.array_equal(g[0, live].astype(np.float32), fx["rho"][live])
        assert np.array_equal(g[1:, live].astype(np.float32), fx["u"][:, live])
        assert np.array_equal(temp.download_fields()[0].astype(np.float32), fx["T"])
    finally:
        temp.close()
        flow.close()


@pytest.mark.parametrize("path", fixtures(), ids=os.path.basename)
@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_coupled_matches_oracle_odd_and_even(path, storage):
    fx = load(path)
    dt = np.float32 if storage == "f32" else np.float64
    flow, temp, cpl = gpu_pair(fx, storage)
    try:
        done = 0
        for k in (1, 2, 7, 13):
            cpl.step(k)
            done += k
            pf, pt, fields, T, _ = oracle_run(fx, dtype=dt, steps=done)
            live = fx["flags"] != VOID
            assert np.array_equal(flow.pops()[:, live], pf[:, live]), f"flow populations after {done} steps"
            assert np.array_equal(temp.pops()[:, live], pt[:, live]), f"temperature populations after {done} steps"
            assert np.array_equal(flow.download_fields()[:, live], fields[:, live].astype(dt).astype(np.float64))
            assert np.array_equal(temp.download_fields()[0], T.astype(dt).astype(np.float64))
    finally:
        temp.close()
        flow.close()


def _cavity(Ni, d):
    nz = Ni + 2 if d == 3 else 1
    dims = (Ni + 2, Ni + 2, nz)
    flags = np.zeros(int(np.prod(dims)), np.uint8)
    O.lib().flbo_cavity_flags(Ni, Ni, Ni if d == 3 else 0, flags.ctypes.data_as(O.C.POINTER(O.C.c_uint8)))
    flags = np.where(flags == O.BBMW, BB, flags).astype(np.uint8)
    f3 = flags.reshape(dims[::-1])
    # hot layer x = 1 and cold layer x = Ni (the whole wall faces, edges included)
    f3[..., 1] = np.where(f3[..., 1] == BB, HOT, f3[..., 1])
    f3[..., Ni] = np.where(f3[..., Ni] == BB, COLD, f3[..., Ni])
    return dims, f3.reshape(-1).copy()


def test_coupled_cavity_48_cubed_two_scalars_matches_oracle():
    """D3Q19 flow with a temperature and a concentration D3Q7 lattice, both
    buoyant (cazsblock2d's two couplings), against the oracle in FP32."""
    dims, flags = _cavity(48, 3)
    N = flags.size
    rng = np.random.default_rng(7)
    T0 = np.where(flags == HOT, 1.5, np.where(flags == COLD, 0.5, 1.0))
    C0 = 1.0 + 0.05 * rng.random(N)
    om_f, om_t, om_c = 1 / 0.6, 1 / 0.7, 1 / 0.9
    # oracle
    flow_o = O.Lattice(19, dims, flags, [(ALL, O.DYN_BGK_SCALAR_FORCE)], [(BB | HOT | COLD, VOID, O.HWBC_NORMAL)],
                       omega=om_f, rho_field=np.ones(N), u_field=np.zeros(3 * N))
    temp_o = O.Lattice(7, dims, flags, [(AABB | BB, O.DYN_SCALAR_BGK), (HOT | COLD, O.DYN_SCALAR_BGK_FIELD)],
                       [(BB, VOID, O.HWBC_NORMAL), (HOT | COLD, VOID, O.HWBC_ANTI)], omega=om_t, rho_field=T0)
    conc_o = O.Lattice(7, dims, flags, [(ALL, O.DYN_SCALAR_BGK)], [(ALL, VOID, O.HWBC_NORMAL)], omega=om_c,
                       rho_field=C0)
    from freelb_amd.lattice import weights
    pf = np.ascontiguousarray(np.repeat(weights(19, np.float32)[:, None], N, axis=1))
    pt = np.ascontiguousarray((weights(7)[:, None] * T0[None, :]).astype(np.float32))
    pc = np.ascontiguousarray((weights(7)[:, None] * C0[None, :]).astype(np.float32))
    gb_t, gb_c = 2e-3, -1e-3
    # GPU
    flow = BlockLattice(19, dims, om_f)
    temp = BlockLattice(7, dims, om_t)
    conc = BlockLattice(7, dims, om_c)
    try:
        for b in (flow, temp, conc):
            b.upload_flags(flags)
        flow.set_tasks([Key_TypePair(ALL, abi.DYN_BGK_SCALAR_FORCE)])
        flow.set_halfway_bc([HalfwayBB(BB | HOT | COLD, VOID)])
        temp.set_tasks([Key_TypePair(AABB | BB, abi.DYN_SCALAR_BGK), Key_TypePair(HOT | COLD, abi.DYN_SCALAR_BGK_FIELD)])
        temp.set_halfway_bc([HalfwayBB(BB, VOID), HalfwayBB(HOT | COLD, VOID, anti=True)])
        conc.set_tasks([Key_TypePair(ALL, abi.DYN_SCALAR_BGK)])
        conc.set_halfway_bc([HalfwayBB(ALL, VOID)])
        temp.set_field_rho(T0)
        conc.set_field_rho(C0)
        temp.share_velocity(flow)
        conc.share_velocity(flow)
        BlockLatManagerCoupling(flow, temp).add_buoyancy(ALL, gb_t, 1.0)
        BlockLatManagerCoupling(flow, conc).add_buoyancy(ALL, gb_c, 1.0)
        flow.upload_pops(pf)
        temp.upload_pops(pt)
        conc.upload_pops(pc)
        done = 0
        for k in (3, 8):
            for _ in range(k):
                flow.step(1)
                temp.step(1)
                conc.step(1)
            done += k
            O.run_coupled(flow_o, pf, [temp_o, conc_o], [pt, pc], [(0, ALL, gb_t, 1.0), (1, ALL, gb_c, 1.0)], k)
            live = flags != VOID
            assert np.array_equal(flow.pops()[:, live], pf[:, live]), f"flow after {done}"
            assert np.array_equal(temp.pops()[:, live], pt[:, live]), f"temperature after {done}"
            assert np.array_equal(conc.pops()[:, live], pc[:, live]), f"concentration after {done}"
    finally:
        for b in (conc, temp, flow):
            b.close()


def test_scalar_sum_conserved_large_2d():
    """1024^2 D2Q9 + D2Q5 closed box in FP64, adiabatic walls, buoyancy from a
    warm blob: the temperature sum stays constant to rounding over 400 steps
    (D2Q5's equilibrium and normal bounce-back conserve it exactly)."""
    Ni = 1022
    dims, flags = _cavity(Ni, 2)
    flags = np.where((flags == HOT) | (flags == COLD), BB, flags).astype(np.uint8)
    N = flags.size
    yy, xx = np.meshgrid(np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
    T0 = (1.0 + 0.5 * np.exp(-((xx - 400.0) ** 2 + (yy - 300.0) ** 2) / 2000.0)).reshape(-1)
    flow = BlockLattice(9, dims, 1 / 0.6, storage="f64")
    temp = BlockLattice(5, dims, 1 / 0.6, storage="f64")
    try:
        for b in (flow, temp):
            b.upload_flags(flags)
        flow.set_tasks([Key_TypePair(AABB | BB, abi.DYN_BGK_SCALAR_FORCE)])
        flow.set_halfway_bc([HalfwayBB(BB, VOID)])
        temp.set_tasks([Key_TypePair(AABB | BB, abi.DYN_SCALAR_BGK)])
        temp.set_halfway_bc([HalfwayBB(BB, VOID)])
        temp.set_field_rho(T0)
        temp.share_velocity(flow)
        BlockLatManagerCoupling(flow, temp).add_buoyancy(AABB | BB, 1e-4, 1.0)
        flow.init_uniform(1.0)
        from freelb_amd.lattice import weights
        temp.upload_pops(weights(5)[:, None] * T0[None, :])
        live = flags != VOID
        s0 = temp.mass_momentum(AABB | BB)[0]
        for _ in range(400):
            flow.step(1)
            temp.step(1)
        s1 = temp.mass_momentum(AABB | BB)[0]
        assert s1 == pytest.approx(s0, rel=1e-12)
        T = temp.download_fields()[0][live]
        assert np.isfinite(T).all() and T.max() < 1.5 and T.min() > 0.99
        u = flow.download_fields()[1:, live]
        assert np.abs(u[1]).max() > 1e-4  # the blob rises
    finally:
        temp.close()
        flow.close()


def test_coupling_argument_checks():
    dims = (10, 9, 8)
    with pytest.raises(FLBError):
        BlockLattice(7, dims, 1.0, storage="f16s")
    with pytest.raises(FLBError):
        BlockLattice(5, dims, 1.0)  # 2-D lattice with nz > 1
    flow = BlockLattice(19, dims, 1.2)
    temp = BlockLattice(7, dims, 1.3)
    other = BlockLattice(7, (10, 9, 9), 1.3)
    temp64 = BlockLattice(7, dims, 1.3, storage="f64")
    try:
        with pytest.raises(FLBError):
            flow.set_tasks([Key_TypePair(AABB, abi.DYN_SCALAR_BGK)])
        with pytest.raises(FLBError):
            temp.set_tasks([Key_TypePair(AABB, abi.DYN_BGK)])
        with pytest.raises(FLBError):
            flow.set_halfway_bc([HalfwayBB(BB, VOID, anti=True)])
        with pytest.raises(FLBError):
            temp.set_halfway_bc([HalfwayBB(BB, VOID, moving=True)])
        with pytest.raises(FLBError):
            other.share_velocity(flow)  # different cells
        with pytest.raises(FLBError):
            temp64.share_velocity(flow)  # different compute type
        with pytest.raises(FLBError):
            flow.share_velocity(temp)  # roles swapped
        cpl = BlockLatManagerCoupling(flow, temp)
        cpl.add_buoyancy(ALL, 1e-3, 1.0)
        cpl.add_buoyancy(ALL, 1e-3, 1.0)
        with pytest.raises(FLBError):
            cpl.add_buoyancy(ALL, 1e-3, 1.0)  # at most two
        cpl.clear()
        cpl.add_buoyancy(ALL, 1e-3, 1.0)
    finally:
        for b in (temp64, other):
            b.close()
    # destroying the scalar block first detaches the flow's coupling; the flow
    # block keeps stepping with no buoyancy (SCALARFORCE = 0)
    flags = np.full(int(np.prod(dims)), AABB, np.uint8)
    flow.upload_flags(np.where(_outer(dims), VOID, flags).astype(np.uint8))
    flow.set_tasks([Key_TypePair(AABB, abi.DYN_BGK_SCALAR_FORCE)])
    flow.init_uniform(1.0)
    temp.share_velocity(flow)
    temp.close()
    flow.step(3)
    p = flow.pops()
    assert np.isfinite(p).all()
    flow.close()


def _outer(dims):
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return ((x == 0) | (x == nx - 1) | (y == 0) | (y == ny - 1) | (z == 0) | (z == nz - 1)).reshape(-1)


def test_coupling_rejects_blocks_on_different_streams():
    """The scalar step reads the flow's VELOCITY field and the flow step the
    scalar's RHO field; on two streams nothing would order them. The streams
    come from the HIP runtime the library itself uses."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    s1, s2 = C.c_void_p(), C.c_void_p()
    assert hip.hipStreamCreate(C.byref(s1)) == 0 and hip.hipStreamCreate(C.byref(s2)) == 0
    dims = (8, 7, 6)
    flow = BlockLattice(19, dims, omega=1.5, stream=s1)
    temp = BlockLattice(7, dims, omega=1.2, stream=s2)
    same = BlockLattice(7, dims, omega=1.2, stream=s1)
    try:
        with pytest.raises(FLBError, match="share one stream"):
            temp.share_velocity(flow)
        with pytest.raises(FLBError, match="share one stream"):
            BlockLatManagerCoupling(flow, temp).add_buoyancy(ALL, 1e-3, 1.0)
        same.share_velocity(flow)
    finally:
        for b in (same, temp, flow):
            b.close()
        hip.hipStreamDestroy(s1)
        hip.hipStreamDestroy(s2)
