"""Small FP64 CPU model of the differentiable step, written with torch so
that autograd supplies a reference gradient for the hand-written GPU adjoint.

Same model as the GPU (BGK with body-force density G, Brinkman drag
-alpha*s*rho*u taken implicitly in u, Guo forcing term as FreeLB's
ForcePopImpl, push streaming with halfway bounce-back at solid cells and at
the ends of non-periodic axes); the objective is the time average over steps
1..T of the x-momentum on the plane x = outlet. Everything is vectorised
over cells, streaming is a scatter through a precomputed permutation.
"""
from __future__ import annotations

import numpy as np
import torch

C19 = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                [1, 1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, -1], [0, 1, 1], [0, -1, -1],
                [1, -1, 0], [-1, 1, 0], [1, 0, -1], [-1, 0, 1], [0, 1, -1], [0, -1, 1]])
W19 = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
OPP = np.array([0] + [i + 1 if i % 2 else i - 1 for i in range(1, 19)])


def push_index(dims, periodic, solid):
    """dst[i, id]: flat index (into [19*N]) that post-collision f*_i(id) streams to."""
    nx, ny, nz = dims
    N = nx * ny * nz
    ids = np.arange(N)
    x, y, z = ids % nx, (ids // nx) % ny, ids // (nx * ny)
    dst = np.empty((19, N), dtype=np.int64)
    for i in range(19):
        wall = np.zeros(N, dtype=bool)
        nb = []
        for v, n, p in ((x + C19[i, 0], nx, periodic[0]), (y + C19[i, 1], ny, periodic[1]),
                        (z + C19[i, 2], nz, periodic[2])):
            out = (v < 0) | (v >= n)
            if p:
                v = v % n
            else:
                wall |= out
                v = np.clip(v, 0, n - 1)
            nb.append(v)
        nid = nb[0] + nx * (nb[1] + ny * nb[2])
        wall |= solid[nid]
        dst[i] = np.where(wall, OPP[i] * N + ids, i * N + nid)
    return dst


class Model:
    def __init__(self, dims, omega, force=(0.0, 0.0, 0.0), alpha=0.0, outlet=None,
                 periodic=(False, False, False), solid=None):
        self.dims = tuple(int(v) for v in dims)
        self.N = int(np.prod(self.dims))
        self.omega = float(omega)
        self.G = torch.tensor(force, dtype=torch.float64)
        self.alpha = float(alpha)
        self.outlet = self.dims[0] - 1 if outlet is None else int(outlet)
        self.solid = np.zeros(self.N, dtype=bool) if solid is None else np.asarray(solid, bool).reshape(-1)
        fluid = ~self.solid
        dst = push_index(self.dims, periodic, self.solid)
        self.fluid = torch.tensor(fluid)
        self.scatter_idx = torch.tensor(dst[:, fluid].reshape(-1))
        self.C = torch.tensor(C19, dtype=torch.float64)
        self.W = torch.tensor(W19, dtype=torch.float64)
        ids = np.arange(self.N)
        self.plane = torch.tensor((ids % self.dims[0] == self.outlet) & fluid)

    def equilibrium(self, rho, u):
        rho = torch.as_tensor(rho, dtype=torch.float64)
        u = torch.as_tensor(u, dtype=torch.float64)
        cu = self.C @ u
        usq = (u * u).sum(0)
        f = self.W[:, None] * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq)
        return f * self.fluid.to(torch.float64)

    def step(self, f, s):
        om = self.omega
        # solid cells hold no populations; rho = 1 there keeps their (unused)
        # values finite, so that no NaN reaches the gradient through a 0 weight
        rho = torch.where(self.fluid, f.sum(0), torch.ones(self.N, dtype=torch.float64))
        j = self.C.T @ f
        D = 1 + self.alpha * s / 2
        u = (j + self.G[:, None] / 2) / (rho * D)
        F = self.G[:, None] - self.alpha * s * rho * u
        cu = self.C @ u
        usq = (u * u).sum(0)
        feq = self.W[:, None] * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq)
        t = 3 * (self.C[:, :, None] - u[None]) + 9 * cu[:, None, :] * self.C[:, :, None]
        Fi = self.W[:, None] * (t * F[None]).sum(1)
        fstar = om * feq + (1 - om) * f + (1 - om / 2) * Fi
        vals = fstar[:, self.fluid].reshape(-1)
        new = torch.zeros(19 * self.N, dtype=torch.float64).scatter(0, self.scatter_idx, vals)
        return new.view(19, self.N)

    def objective_term(self, f):
        return (self.C[:, 0][:, None] * f[:, self.plane]).sum()

    def run(self, f0, s, T):
        """J and the final state after T steps."""
        f = f0
        J = torch.zeros((), dtype=torch.float64)
        for _ in range(T):
            f = self.step(f, s)
            J = J + self.objective_term(f)
        return J / T, f
