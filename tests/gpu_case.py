"""Hands an oracle case (oracle.Lattice + initial pops) to the GPU path."""
from __future__ import annotations

import numpy as np

from freelb_amd import BlockLattice, HalfwayBB, Key_TypePair


def gpu_lattice(case, storage="f32", pops=None):
    """BlockLattice on the GPU with the same flags, tasks, BCs and fields as `case`."""
    lat = BlockLattice(case.q, (case.nx, case.ny, case.nz), omega=case.omega, storage=storage,
                       overlap=int(case.L.overlap))
    lat.upload_flags(case.flags)
    lat.set_tasks([Key_TypePair(t[0], t[1], float(t[2]) if len(t) > 2 else 0.0) for t in case.tasks])
    lat.set_halfway_bc([HalfwayBB(b[0], b[1], b[2] == 1) for b in case.bcs])
    if case.rho_field is not None:
        lat.set_field_rho(case.rho_field)
    if case.u_field is not None:
        lat.set_field_u(case.u_field)
    if case.mrt_rates is not None:
        lat.set_mrt_rates(case.mrt_rates)
    if any(getattr(case, "force", ())):
        lat.set_force(case.force[:(3 if case.q == 19 else 2)])
    if pops is not None:
        lat.upload_pops(pops)
    return lat


def interior_mask(case, overlap=1):
    """Cells the kernels visit (the overlap layer excluded)."""
    m = np.zeros((case.nz, case.ny, case.nx), dtype=bool)
    if case.nz > 1:
        m[overlap:-overlap, overlap:-overlap, overlap:-overlap] = True
    else:
        m[:, overlap:-overlap, overlap:-overlap] = True
    return m.reshape(-1)
