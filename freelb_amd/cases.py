"""Lid-driven cavity set-up as FreeLB's cavity drivers build it.

Flags follow benchmarks/cavity3d/cavity3d.cpp:121-142 with the AABB rule of
geometry/basic_geometry.hh:38-48: a void shell of one cell (1), walls on the
outer layer of the n^3 cavity (4), the lid on the top layer without its edges
(8), fluid inside (2). In the halfway mode every non-void cell collides and
the walls sit at the half link to the void shell (bounceback::normal for 4,
bounceback::movingwall for 8, boundary/bounce_back_boundary.h:169-256); the
lid's RHO / VELOCITY fields hold rho = 1, u = (u_lid, 0, 0).
"""
from __future__ import annotations

import numpy as np

from . import abi
from .lattice import BlockLattice, HalfwayBB, Key_TypePair

VOID, FLUID, WALL, LID = 1, 2, 4, 8


def cavity_flags(n: int, ny: int | None = None, nz: int | None = None) -> np.ndarray:
    """Flags of an n x ny x nz cavity plus the void shell, FreeLB index order."""
    ny = n if ny is None else ny
    nz = n if nz is None else nz
    f = np.full((nz + 2, ny + 2, n + 2), VOID, dtype=np.uint8)
    f[1:-1, 1:-1, 1:-1] = WALL
    f[2:-2, 2:-2, 2:-2] = FLUID
    f[nz, 2:-2, 2:-2] = LID
    return f.reshape(-1)


def cavity_block(n, storage: str = "f32", collision: str = "bgk", tau: float = 0.53,
                 u_lid: float = 0.1, cs: float = 0.17, u0_amp: float = 0.0, seed: int = 0,
                 mrt_rates=None, stream=None) -> BlockLattice:
    """A D3Q19 halfway cavity of n^3 (or n = (nx, ny, nz)) cells on the GPU, at equilibrium with
    rho = 1 and u = 0, or u drawn i.i.d. from U(-u0_amp, u0_amp) per
    component on every non-void cell (numpy default_rng(seed)).

    collision: "bgk" (collision::BGK) or "mrt_les" (MRT whose shear rates
    take the Smagorinsky omega with constant cs)."""
    omega = float(np.float32(1.0) / np.float32(tau))
    nx, ny, nz = (n, n, n) if np.isscalar(n) else (int(v) for v in n)
    lat = BlockLattice(19, (nx + 2, ny + 2, nz + 2), omega=omega, storage=storage, stream=stream)
    try:
        flags = cavity_flags(nx, ny, nz)
        lat.upload_flags(flags)
        mask = FLUID | WALL | LID
        if collision == "bgk":
            lat.set_tasks([Key_TypePair(mask, abi.DYN_BGK)])
        elif collision == "mrt":
            lat.set_tasks([Key_TypePair(mask, abi.DYN_MRT)])
            lat.set_mrt_rates(mrt_rates)
        elif collision == "mrt_les":
            lat.set_tasks([Key_TypePair(mask, abi.DYN_SMAG_MRT, cs)])
            lat.set_mrt_rates(mrt_rates)
        else:
            raise ValueError(f"unknown collision {collision!r}")
        lat.set_halfway_bc([HalfwayBB(WALL, VOID, False), HalfwayBB(LID, VOID, True)])
        N = flags.size
        u = np.zeros((3, N))
        u[0, flags == LID] = u_lid
        lat.set_field_rho(np.ones(N))
        lat.set_field_u(u)
        # f = feq(1, u0) (with u0 = 0 this is exactly w_i, FreeLB's initial state)
        u0 = np.zeros((3, N))
        if u0_amp:
            rng = np.random.default_rng(seed)
            live = flags != VOID
            u0[:, live] = rng.uniform(-u0_amp, u0_amp, size=(3, int(live.sum())))
        lat.init_equilibrium(np.ones(N), u0)
    except Exception:
        lat.close()
        raise
    return lat


def solid_fraction(dims, seed: int, sigma: float = 4.0) -> np.ndarray:
    """Brinkman solid fraction s in (0, 1) per cell, FreeLB index order: a
    seeded N(0,1) field (numpy default_rng(seed)), a periodic Gaussian filter
    of width sigma cells (in Fourier space), rescaled to zero mean and unit
    variance, then the logistic sigmoid."""
    nx, ny, nz = (int(v) for v in dims)
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((nz, ny, nx))
    kz, ky, kx = np.meshgrid(np.fft.fftfreq(nz), np.fft.fftfreq(ny), np.fft.fftfreq(nx), indexing="ij")
    filt = np.exp(-2.0 * (np.pi * sigma) ** 2 * (kx ** 2 + ky ** 2 + kz ** 2))
    sm = np.real(np.fft.ifftn(np.fft.fftn(g) * filt))
    sm = (sm - sm.mean()) / sm.std()
    return (1.0 / (1.0 + np.exp(-sm))).reshape(-1)


def channel_adjoint(dims=(192, 96, 96), storage: str = "f32", tau: float = 0.6, g: float = 1e-6,
                    alpha: float = 0.1, seed: int = 7, sigma: float = 4.0):
    """The differentiable channel: periodic in x, halfway walls beyond y and z,
    body force g along x, Brinkman drag alpha*s with s from solid_fraction,
    objective on the last x plane, at rest (rho = 1, u = 0) at t = 0."""
    from .adjoint import Adjoint
    adj = Adjoint(dims, omega=1.0 / tau, force=(g, 0.0, 0.0), brinkman_alpha=alpha,
                  outlet_x=int(dims[0]) - 1, periodic=(True, False, False), storage=storage)
    try:
        N = int(np.prod(dims))
        s = solid_fraction(dims, seed, sigma)
        adj.set_fraction(s)
        adj.init_equilibrium(np.ones(N), np.zeros((3, N)))
    except Exception:
        adj.close()
        raise
    return adj, s


def channel_dims(n: int = 64, lo=(64, 32, 32), hi=(128, 64, 64), seed: int = 11):
    """n seeded channel extents, each axis uniform in [lo, hi] (numpy
    default_rng(seed), integers drawn axis by axis, box by box)."""
    rng = np.random.default_rng(seed)
    return [tuple(int(rng.integers(a, b + 1)) for a, b in zip(lo, hi)) for _ in range(n)]


def channel_batch(dims_list, storage: str = "f32", tau: float = 0.6, g: float = 1e-6,
                  alpha: float = 0.1, seed: int = 11, sigma: float = 4.0):
    """One channel_adjoint box per extent in dims_list (box k draws its solid
    fraction with seed 1000*seed + k) and the AdjointBatch over all of them.
    Returns (batch, members, fractions); closing is the caller's."""
    from .adjoint import AdjointBatch
    members, fields = [], []
    try:
        for k, dims in enumerate(dims_list):
            adj, s = channel_adjoint(dims, storage, tau, g, alpha, seed=1000 * seed + k, sigma=sigma)
            members.append(adj)
            fields.append(s)
        batch = AdjointBatch(members)
    except Exception:
        for m in members:
            m.close()
        raise
    return batch, members, fields


COVERED = 16  # coarse cells under a finer patch beyond its outer coarse layer


def refined_cavity(n: int = 128, patch: int = 64, storage: str = "f32", tau: float = 0.6, u_lid: float = 0.05,
                   stream=None):
    """The two-level lid-driven cavity of examples/cavref2d/cavref2d.cpp in 3D:
    an n^3 D3Q19 coarse cavity (halfway walls, moving lid) with a centred
    patch of patch^3 fine cells (patch / 2 coarse cells per axis) at twice the
    resolution. Fluid cells collide with BGK writing their moments to the
    fields (rhoU<CELL, true>), wall cells with BGK<useFieldrhoU>; the fine
    omega is RefineConverter::getOmegaF(omega) = 1 / (2 / omega - 1/2) in
    float (lbm/unit_converter.h:627). Returns (coarse, fine, Refinement)."""
    from .lattice import Refinement
    if patch % 2 or patch < 4 or patch // 2 + 4 > n:
        raise ValueError("patch must be even, >= 4 and leave room for the coarse stencils")
    m = patch // 2
    origin = 1 + (n - m) // 2
    omega_c = np.float32(1.0) / np.float32(tau)
    omega_f = np.float32(1) / (np.float32(2) / omega_c - np.float32(0.5))
    flags = cavity_flags(n).reshape(n + 2, n + 2, n + 2)
    flags[origin + 1:origin + m - 1, origin + 1:origin + m - 1, origin + 1:origin + m - 1] = COVERED
    flags = flags.reshape(-1)
    tasks = [Key_TypePair(FLUID, abi.DYN_BGK_WRITE), Key_TypePair(WALL | LID, abi.DYN_BGK_FIELD)]
    coarse = BlockLattice(19, (n + 2,) * 3, omega=float(omega_c), storage=storage, stream=stream)
    fine = None
    try:
        N = flags.size
        coarse.upload_flags(flags)
        coarse.set_tasks(tasks)
        coarse.set_halfway_bc([HalfwayBB(WALL, VOID, False), HalfwayBB(LID, VOID, True)])
        u = np.zeros((3, N))
        u[0, flags == LID] = float(np.float32(u_lid))
        coarse.set_field_rho(np.ones(N))
        coarse.set_field_u(u)
        coarse.init_equilibrium(np.ones(N), np.zeros((3, N)))
        fn = 2 * m + 4
        fine = BlockLattice(19, (fn,) * 3, omega=float(omega_f), storage=storage, stream=stream)
        ff = np.full((fn, fn, fn), COVERED, dtype=np.uint8)
        ff[1:-1, 1:-1, 1:-1] = FLUID
        NF = ff.size
        fine.upload_flags(ff.reshape(-1))
        fine.set_tasks(tasks)
        fine.set_field_rho(np.ones(NF))
        fine.set_field_u(np.zeros((3, NF)))
        fine.init_equilibrium(np.ones(NF), np.zeros((3, NF)))
        ref = Refinement(coarse, fine, (origin,) * 3, stream=stream)
    except Exception:
        coarse.close()
        if fine is not None:
            fine.close()
        raise
    return coarse, fine, ref


# Flags of the differentially heated cavity: walls (4) adiabatic, the x = 1
# layer hot (8) and the x = n layer cold (16), edges included.
HOT, COLD = 8, 16


def heated_cavity_flags(n: int, d: int = 3) -> np.ndarray:
    if d == 3:
        f = np.full((n + 2, n + 2, n + 2), VOID, dtype=np.uint8)
        f[1:-1, 1:-1, 1:-1] = WALL
        f[2:-2, 2:-2, 2:-2] = FLUID
        f[1:-1, 1:-1, 1] = HOT
        f[1:-1, 1:-1, n] = COLD
    else:
        f = np.full((1, n + 2, n + 2), VOID, dtype=np.uint8)
        f[0, 1:-1, 1:-1] = WALL
        f[0, 2:-2, 2:-2] = FLUID
        f[0, 1:-1, 1] = HOT
        f[0, 1:-1, n] = COLD
    return f.reshape(-1)


def heated_cavity(n: int = 256, d: int = 3, storage: str = "f32", tau_flow: float = 0.6, tau_temp: float = 0.7,
                  gbeta: float = 1e-4, t_hot: float = 1.5, t_cold: float = 0.5, t_init: float = 1.0):
    """A differentially heated cavity of n^d cells on the GPU: a flow block
    (D3Q19 / D2Q9, BGKForce with the scalar buoyancy force, halfway walls) and
    a temperature block (D3Q7 / D2Q5, advected by the flow's VELOCITY field,
    adiabatic walls and fixed-temperature hot / cold walls), coupled by
    force::Buoyancy; FreeLB's coupled-lattice loop of
    examples/cazsblock2d/cazsblock2d.cpp:385-409 without the cellular
    automaton. Returns (flow, temperature, coupling)."""
    from .lattice import BlockLatManagerCoupling, weights
    qf, qt = (19, 7) if d == 3 else (9, 5)
    dims = (n + 2, n + 2, n + 2) if d == 3 else (n + 2, n + 2, 1)
    flags = heated_cavity_flags(n, d)
    alls = FLUID | WALL | HOT | COLD
    flow = BlockLattice(qf, dims, 1.0 / tau_flow, storage=storage)
    temp = BlockLattice(qt, dims, 1.0 / tau_temp, storage=storage)
    for b in (flow, temp):
        b.upload_flags(flags)
    flow.set_tasks([Key_TypePair(alls, abi.DYN_BGK_SCALAR_FORCE)])
    flow.set_halfway_bc([HalfwayBB(WALL | HOT | COLD, VOID)])
    temp.set_tasks([Key_TypePair(FLUID | WALL, abi.DYN_SCALAR_BGK),
                    Key_TypePair(HOT | COLD, abi.DYN_SCALAR_BGK_FIELD)])
    temp.set_halfway_bc([HalfwayBB(WALL, VOID), HalfwayBB(HOT | COLD, VOID, anti=True)])
    T0 = np.where(flags == HOT, t_hot, np.where(flags == COLD, t_cold, t_init)).astype(np.float64)
    temp.set_field_rho(T0)
    temp.share_velocity(flow)
    cpl = BlockLatManagerCoupling(flow, temp)
    cpl.add_buoyancy(alls, gbeta, t_init)
    flow.init_uniform(1.0)
    real = np.float64 if storage == "f64" else np.float32
    temp.upload_pops((weights(qt, real)[:, None] * T0.astype(real)[None, :]).astype(real))
    return flow, temp, cpl
