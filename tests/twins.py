"""Oracle twins of the GPU blocks that freelb_amd.cases builds, so that batch,
halo and workload tests compare the HIP path with the oracle rather than with
other GPU runs.

A twin is an oracle.Lattice with the block's flags, task table, halfway
managers, fields and omega; it starts from the populations the GPU block
holds (downloaded before the first step), so both sides begin from the same
bits.

A periodic axis has no counterpart in FreeLB's periodic-shift streaming (it
wraps a whole direction array, row into row). FreeLB closes a periodic
domain with FixedPeriodicBoundary (boundary/periodic_boundary.h:28-120),
whose virtual cells beyond the boundary mirror the real cells on the other
side. The twin of a block periodic in x does the same: one virtual column on
each side of x, refreshed from the opposite real column before every step
and colliding like it, so that the populations streaming out of one face
enter through the other.
"""
from __future__ import annotations

import numpy as np

from freelb_amd import cases
from oracle import oracle as O


def cavity_twin(shape, collision="bgk", tau=0.53, u_lid=0.1, cs=0.17, rates=None):
    """Twin of cases.cavity_block(shape, collision=..., tau, u_lid, cs, mrt_rates=rates)."""
    nx, ny, nz = (shape, shape, shape) if np.isscalar(shape) else (int(v) for v in shape)
    flags = cases.cavity_flags(nx, ny, nz)
    mask = cases.FLUID | cases.WALL | cases.LID
    if collision == "bgk":
        tasks = [(mask, O.DYN_BGK)]
    elif collision == "mrt":
        tasks = [(mask, O.DYN_MRT)]
    elif collision == "mrt_les":
        tasks = [(mask, O.DYN_SMAG_MRT, float(np.float32(cs)))]  # the task's C_s is a float
    else:
        raise ValueError(collision)
    bcs = [(cases.WALL, cases.VOID, O.HWBC_NORMAL), (cases.LID, cases.VOID, O.HWBC_MOVING)]
    N = flags.size
    u = np.zeros((3, N))
    u[0, flags == cases.LID] = u_lid
    omega = float(np.float32(1.0) / np.float32(tau))
    return O.Lattice(19, (nx + 2, ny + 2, nz + 2), flags, tasks, bcs, omega=omega, rho_field=np.ones(N),
                     u_field=u, mrt_rates=rates)


def interior(dims, overlap=1, periodic=(False, False, False)):
    """Cells a block updates (its overlap layer excluded on non-periodic axes)."""
    nx, ny, nz = dims
    m = np.zeros((nz, ny, nx), dtype=bool)
    sl = [slice(None) if periodic[a] or n == 1 else slice(overlap, n - overlap)
          for a, n in zip((2, 1, 0), (nz, ny, nx))]
    m[tuple(sl)] = True
    return m.reshape(-1)


class PeriodicXTwin:
    """Oracle twin of a block periodic in x (nx cells, no overlap in x): an
    oracle block of nx + 2 columns whose columns 0 and nx + 1 mirror the real
    columns nx and 1 (FixedPeriodicBoundary's virtual cells)."""

    def __init__(self, q, dims, flags, tasks, bcs, omega, rho_field=None, u_field=None, rates=None):
        nx, ny, nz = dims
        self.q, self.dims = q, dims
        d = 3 if q == 19 else 2
        f = np.asarray(flags, np.uint8).reshape(nz, ny, nx)
        self.flags = np.concatenate([f[:, :, -1:], f, f[:, :, :1]], axis=2).reshape(-1)
        self.N = self.flags.size

        def pad(a, comps):
            if a is None:
                return None
            a = np.asarray(a, np.float64).reshape(comps, nz, ny, nx)
            return np.concatenate([a[..., -1:], a, a[..., :1]], axis=3).reshape(comps, -1)

        rf = pad(rho_field, 1)
        uf = pad(u_field, d)
        self.lat = O.Lattice(q, (nx + 2, ny, nz), self.flags, tasks, bcs, omega=omega,
                             rho_field=None if rf is None else rf[0], u_field=uf, mrt_rates=rates)

    def widen(self, pops):
        """[q][nx ny nz] of the periodic block -> the twin's [q][(nx + 2) ny nz]."""
        nx, ny, nz = self.dims
        p = np.asarray(pops).reshape(self.q, nz, ny, nx)
        return np.ascontiguousarray(np.concatenate([p[..., -1:], p, p[..., :1]], axis=3).reshape(self.q, -1))

    def narrow(self, pops):
        nx, ny, nz = self.dims
        return np.ascontiguousarray(pops.reshape(self.q, nz, ny, nx + 2)[..., 1:-1].reshape(self.q, -1))

    def run(self, pops, steps):
        """Advance widened pops in place: refresh the virtual columns, one step, repeat."""
        nx, ny, nz = self.dims
        v = pops.reshape(self.q, nz, ny, nx + 2)
        for _ in range(steps):
            v[..., 0] = v[..., nx]
            v[..., nx + 1] = v[..., 1]
            self.lat.run(pops, 1)
        return pops
