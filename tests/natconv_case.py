"""The coupled flow + temperature case of the natconv fixtures (test helper).

A differentially heated cavity assembled the way FreeLB's coupled examples
assemble their lattices (examples/cazsblock2d/cazsblock2d.cpp:225-409): a
flow lattice (D2Q9 / D3Q19) whose cells collide with BGKForce<forcerhoU<
ScalarForce, true>>, a temperature lattice (D2Q5 / D3Q7) advected by the
flow's VELOCITY field, and force::Buoyancy coupling the two. The fixture holds
FreeLB's own run (oracle/_ref/natconv_ref); the oracle and the GPU path are
set up here from the same arrays.
"""
from __future__ import annotations

import glob
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

VOID, AABB, BB, HOT, COLD = 1, 2, 4, 8, 16
ALL = AABB | BB | HOT | COLD


def fixtures():
    return sorted(glob.glob(os.path.join(GOLDEN, "natconv_*.npz")))


def load(path):
    z = np.load(path)
    return {k: z[k] for k in z.files}


def params(fx):
    om_f, om_t, gbeta, th, tc, t0 = (float(v) for v in fx["params"])
    return dict(omega_flow=om_f, omega_temp=om_t, gbeta=gbeta, T_hot=th, T_cold=tc, T_init=t0)


def oracle_run(fx, dtype=np.float32, steps=None):
    """Runs the oracle's coupled loop from the fixture's initial state; returns
    (flow pops, temperature pops, flow fields [(d+1)][N], T field, SCALARFORCE)."""
    import oracle.oracle as O
    p = params(fx)
    dims = tuple(int(v) for v in fx["dims"])
    N = int(np.prod(dims))
    d, qf, qt = int(fx["d"]), int(fx["qf"]), int(fx["qt"])
    flow = O.Lattice(qf, dims, fx["flags"], [(ALL, O.DYN_BGK_SCALAR_FORCE)], [(BB | HOT | COLD, VOID, O.HWBC_NORMAL)],
                     omega=p["omega_flow"], rho_field=np.ones(N), u_field=np.zeros(d * N))
    temp = O.Lattice(qt, dims, fx["flags"],
                     [(AABB | BB, O.DYN_SCALAR_BGK), (HOT | COLD, O.DYN_SCALAR_BGK_FIELD)],
                     [(BB, VOID, O.HWBC_NORMAL), (HOT | COLD, VOID, O.HWBC_ANTI)],
                     omega=p["omega_temp"], rho_field=fx["T0"].astype(np.float64))
    pf = np.ascontiguousarray(fx["flow0"].astype(dtype))
    pt = np.ascontiguousarray(fx["temp0"].astype(dtype))
    n = int(fx["steps"]) if steps is None else int(steps)
    O.run_coupled(flow, pf, [temp], [pt], [(0, ALL, p["gbeta"], p["T_init"])], n)
    return pf, pt, flow.fields(), temp.buffers["rho_field"].copy(), flow.buffers["sforce"].copy()


def gpu_pair(fx, storage="f32"):
    """The same case on the GPU: (flow block, temperature block, coupling)."""
    from freelb_amd import BlockLatManagerCoupling, BlockLattice, HalfwayBB, Key_TypePair, abi
    p = params(fx)
    dims = tuple(int(v) for v in fx["dims"])
    qf, qt = int(fx["qf"]), int(fx["qt"])
    real = np.float64 if storage == "f64" else np.float32
    flow = BlockLattice(qf, dims, p["omega_flow"], storage=storage)
    temp = BlockLattice(qt, dims, p["omega_temp"], storage=storage)
    for b in (flow, temp):
        b.upload_flags(fx["flags"])
    flow.set_tasks([Key_TypePair(ALL, abi.DYN_BGK_SCALAR_FORCE)])
    flow.set_halfway_bc([HalfwayBB(BB | HOT | COLD, VOID)])
    temp.set_tasks([Key_TypePair(AABB | BB, abi.DYN_SCALAR_BGK), Key_TypePair(HOT | COLD, abi.DYN_SCALAR_BGK_FIELD)])
    temp.set_halfway_bc([HalfwayBB(BB, VOID), HalfwayBB(HOT | COLD, VOID, anti=True)])
    temp.set_field_rho(fx["T0"].astype(np.float64))
    temp.share_velocity(flow)
    cpl = BlockLatManagerCoupling(flow, temp)
    cpl.add_buoyancy(ALL, p["gbeta"], p["T_init"])
    flow.upload_pops(fx["flow0"].astype(real))
    temp.upload_pops(fx["temp0"].astype(real))
    return flow, temp, cpl
