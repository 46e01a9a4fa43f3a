"""Host-side mirror of FreeLB's block-lattice interface over the C ABI.

``BlockLattice`` is one FreeLB ``BlockLattice`` resident on the GPU. Its
methods keep the reference's call shape and names:

    lat.ApplyCellDynamics()   # BlockLattice::ApplyCellDynamics<TaskSelector>(flags)
    lat.Stream()              # BlockLattice::Stream
    lat.ApplyBoundaries()     # BlockBoundaryManager::Apply (fused; order check only)

(data_struct/block_lattice.hh:64-92, boundary/basic_boundary.h:345-347). The
dynamics of a cell is chosen by a first-match task table, as FreeLB's
``TaskSelector<uint8_t, CELL, Key_TypePair<mask, Dynamics>...>`` does
(utils/tmp.h:283-309); halfway bounce-back managers are registered with
``set_halfway_bc`` (boundary/boundary.h:75-129).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Iterable, Sequence

import numpy as np

from . import abi


@dataclass(frozen=True)
class Key_TypePair:
    """FreeLB's tmp::Key_TypePair<mask, Dynamics> (utils/tmp.h:194-202)."""
    mask: int
    dynamics: int
    smagorinsky: float = 0.0


@dataclass(frozen=True)
class HalfwayBB:
    """One BBLikeFixedBlockBdManager<bounceback::normal|movingwall|anti_simple>
    (boundary/boundary.h:75-129, boundary/bounce_back_boundary.h:169-256).
    `anti` (D2Q5 / D3Q7 blocks) is the fixed-value wall anti_simple, whose
    value is the cell's RHO (scalar) field."""
    cell_mask: int
    void_mask: int = 1
    moving: bool = False
    anti: bool = False


def weights(q: int, dtype=np.float64) -> np.ndarray:
    """Lattice weights as FreeLB forms them, T(1)/T(18) etc. (lbm/lattice_set.h)."""
    one = dtype(1)
    if q == 19:
        return np.array([one / dtype(3)] + [one / dtype(18)] * 6 + [one / dtype(36)] * 12, dtype=dtype)
    if q == 9:
        return np.array([dtype(4) / dtype(9)] + [one / dtype(9)] * 4 + [one / dtype(36)] * 4, dtype=dtype)
    if q == 5:
        return np.array([one / dtype(3)] + [one / dtype(6)] * 4, dtype=dtype)
    if q == 7:
        return np.array([one / dtype(4)] + [one / dtype(8)] * 6, dtype=dtype)
    raise ValueError(f"unknown lattice q={q}")


def dimension(q: int) -> int:
    return 3 if q in (19, 7) else 2


class BlockLattice:
    def __init__(self, q: int, dims: Sequence[int], omega: float, storage: str = "f32",
                 overlap: int = 1, periodic: Sequence[bool] = (False, False, False), stream=None):
        L = abi.lib()
        self.q = int(q)
        self.d = dimension(self.q)
        nx, ny = int(dims[0]), int(dims[1])
        nz = int(dims[2]) if len(dims) > 2 else 1
        self.dims = (nx, ny, nz)
        self.N = nx * ny * nz
        self.storage = storage
        desc = abi.BlockDesc()
        desc.q = self.q
        desc.storage = abi.STORAGE[storage]
        desc.nx, desc.ny, desc.nz = nx, ny, nz
        desc.overlap = int(overlap)
        for a in range(3):
            desc.periodic[a] = 1 if (a < len(periodic) and periodic[a]) else 0
        desc.omega = float(omega)
        desc.stream = stream
        h = C.c_void_p()
        abi.check(L.flb_block_create(C.byref(desc), C.byref(h)))
        self._h = h
        self._real = np.float64 if storage == "f64" else np.float32

    # ---- lifecycle -------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            abi.check(abi.lib().flb_block_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def info(self) -> abi.BlockInfo:
        inf = abi.BlockInfo()
        abi.check(abi.lib().flb_get_info(self._h, C.byref(inf)))
        return inf

    def synchronize(self) -> None:
        abi.check(abi.lib().flb_synchronize(self._h))

    # ---- set-up ------------------------------------------------------------
    def set_tasks(self, tasks: Iterable) -> None:
        tasks = [t if isinstance(t, Key_TypePair) else Key_TypePair(*t) for t in tasks]
        arr = (abi.Task * max(1, len(tasks)))()
        for k, t in enumerate(tasks):
            arr[k].mask, arr[k].dynamics, arr[k].smagorinsky = t.mask, t.dynamics, t.smagorinsky
        abi.check(abi.lib().flb_set_tasks(self._h, arr, len(tasks)))

    def set_halfway_bc(self, bcs: Iterable) -> None:
        bcs = [b if isinstance(b, HalfwayBB) else HalfwayBB(*b) for b in bcs]
        arr = (abi.HWBC * max(1, len(bcs)))()
        for k, b in enumerate(bcs):
            arr[k].cell_mask, arr[k].void_mask = b.cell_mask, b.void_mask
            arr[k].kind = abi.HWBC_ANTI if b.anti else (abi.HWBC_MOVING if b.moving else abi.HWBC_NORMAL)
        abi.check(abi.lib().flb_set_halfway_bc(self._h, arr, len(bcs)))

    def set_mrt_rates(self, rates=None) -> None:
        if rates is None:
            abi.check(abi.lib().flb_set_mrt_rates(self._h, None, self.q))
            return
        r = np.ascontiguousarray(rates, dtype=np.float64)
        abi.check(abi.lib().flb_set_mrt_rates(self._h, abi.ptr(r, C.c_double), r.size))

    def set_ghost_overlap(self, enable: bool = True) -> None:
        """The overlap layer mirrors neighbouring blocks (see BlockHalo)."""
        abi.check(abi.lib().flb_set_ghost_overlap(self._h, int(bool(enable))))

    def set_force(self, force) -> None:
        """CONSTFORCE of the BGKForce cells (force::ConstForce, lbm/force.h:86-107)."""
        F = np.ascontiguousarray(np.asarray(force, dtype=np.float64).reshape(-1))
        abi.check(abi.lib().flb_set_force(self._h, abi.ptr(F, C.c_double), int(F.size)))

    def share_velocity(self, flow: "BlockLattice | None") -> None:
        """This (D2Q5 / D3Q7) block is advected by flow's VELOCITY field: the
        field reference FreeLB passes to a temperature / concentration
        lattice's constructor (examples/cazsblock2d/cazsblock2d.cpp:240-251).
        None detaches."""
        abi.check(abi.lib().flb_share_velocity(self._h, flow.handle if flow is not None else None))
        self._vel_src = flow  # keep the flow block alive while it is referenced

    def set_omega(self, omega: float) -> None:
        abi.check(abi.lib().flb_set_omega(self._h, float(omega)))

    def upload_flags(self, flags: np.ndarray) -> None:
        f = np.ascontiguousarray(flags, dtype=np.uint8).reshape(-1)
        if f.size != self.N:
            raise ValueError(f"flags has {f.size} cells, block has {self.N}")
        abi.check(abi.lib().flb_upload_flags(self._h, abi.ptr(f, C.c_uint8)))

    def set_field_rho(self, rho: np.ndarray) -> None:
        r = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        if r.size != self.N:
            raise ValueError("rho field size mismatch")
        abi.check(abi.lib().flb_set_field_rho(self._h, abi.ptr(r, C.c_double)))

    def set_field_u(self, u: np.ndarray) -> None:
        v = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        if v.size != self.d * self.N:
            raise ValueError("velocity field size mismatch (expects d*N SoA)")
        abi.check(abi.lib().flb_set_field_u(self._h, abi.ptr(v, C.c_double)))

    def download_fields(self) -> np.ndarray:
        """RHO and VELOCITY fields, [(d+1)][N] (what BGK_WRITE cells stored)."""
        out = np.empty((self.d + 1, self.N), dtype=np.float64)
        abi.check(abi.lib().flb_download_fields(self._h, abi.ptr(out, C.c_double)))
        return out

    def store_moments(self, mask: int = 0xFE) -> None:
        """rhoU<CELL, true> read-out: rho, u of cells with flag & mask go to the fields."""
        abi.check(abi.lib().flb_store_moments(self._h, int(mask) & 0xFF))

    def enable_tolerance(self) -> None:
        """EnableToleranceRho / EnableToleranceU: the current fields become the old ones."""
        abi.check(abi.lib().flb_enable_tolerance(self._h))

    def tolerance(self, which: str = "u", shift: int = 0) -> float:
        """getToleranceU / getTolU(shift) (which="u") or the RHO form (which="rho"):
        max |field - old| over the cells [shift, n - shift), then old = field."""
        r = C.c_double()
        abi.check(abi.lib().flb_tolerance(self._h, 0 if which == "rho" else 1, int(shift), C.byref(r)))
        return r.value

    # ---- populations -------------------------------------------------------
    def upload_pops(self, pops: np.ndarray) -> None:
        p = np.ascontiguousarray(pops, dtype=self._real).reshape(-1)
        if p.size != self.q * self.N:
            raise ValueError("population array size mismatch (expects q*N SoA)")
        if self._real is np.float64:
            abi.check(abi.lib().flb_upload_pops_f64(self._h, abi.ptr(p, C.c_double)))
        else:
            abi.check(abi.lib().flb_upload_pops_f32(self._h, abi.ptr(p, C.c_float)))

    def init_uniform(self, rho0: float = 1.0) -> None:
        """f_i = rho0 * w_i everywhere (BlockLatticeManager ctor, block_lattice.hh:676-681)."""
        w = weights(self.q, self._real)
        pops = np.repeat((self._real(rho0) * w)[:, None], self.N, axis=1)
        self.upload_pops(pops)

    def init_equilibrium(self, rho: np.ndarray, u: np.ndarray) -> None:
        rhou = np.concatenate([np.asarray(rho, dtype=np.float64).reshape(1, -1),
                               np.asarray(u, dtype=np.float64).reshape(self.d, -1)], axis=0)
        rhou = np.ascontiguousarray(rhou)
        abi.check(abi.lib().flb_init_equilibrium(self._h, abi.ptr(rhou, C.c_double)))

    def pops(self) -> np.ndarray:
        out = np.empty((self.q, self.N), dtype=self._real)
        if self._real is np.float64:
            abi.check(abi.lib().flb_download_pops_f64(self._h, abi.ptr(out, C.c_double)))
        else:
            abi.check(abi.lib().flb_download_pops_f32(self._h, abi.ptr(out, C.c_float)))
        return out

    # ---- the hot path -------------------------------------------------------
    def ApplyCellDynamics(self, count: int = 0) -> None:
        """The fused collide-and-stream step, gated by the block's level
        (BlockLatticeManager::ApplyCellDynamics(count, ...): the block steps
        when count % 2^(max_level - level) == 0)."""
        abi.check(abi.lib().flb_apply_cell_dynamics(self._h, int(count)))

    def Stream(self, count: int = 0) -> None:
        abi.check(abi.lib().flb_stream(self._h, int(count)))

    def set_level(self, level: int, max_level: int) -> None:
        """Refinement level of the block and the finest level of its hierarchy."""
        abi.check(abi.lib().flb_set_level(self._h, int(level), int(max_level)))
        self.level, self.max_level = int(level), int(max_level)

    def ApplyBoundaries(self) -> None:
        abi.check(abi.lib().flb_apply_halfway_bc(self._h))

    def step(self, n: int = 1) -> None:
        abi.check(abi.lib().flb_step(self._h, int(n)))

    def set_watchdog(self, every: int) -> None:
        """step() checks the FP64 mass / momentum every `every` steps and
        raises FLBError (FLB_ERR_NONFINITE) on a NaN / Inf; 0 switches it off."""
        abi.check(abi.lib().flb_set_watchdog(self._h, int(every)))

    # ---- read-out -------------------------------------------------------------
    def moments(self, mask: int = 0xFE) -> np.ndarray:
        """rho, u of cells with flag & mask (moment::rhoU<CELL, true>); [d+1][N]."""
        out = np.empty((self.d + 1, self.N), dtype=self._real)
        abi.check(abi.lib().flb_moments(self._h, mask & 0xFF, out.ctypes.data_as(C.c_void_p)))
        return out

    def mass_momentum(self, mask: int = 0xFE) -> np.ndarray:
        out = np.zeros(4, dtype=np.float64)
        abi.check(abi.lib().flb_reduce_mass_momentum(self._h, mask & 0xFF, abi.ptr(out, C.c_double)))
        return out


class BlockLatManagerCoupling:
    """FreeLB's BlockLatManagerCoupling(flow, scalar) (data_struct/block_lattice.h:436-649)
    running force::Buoyancy<NSCELL, CELL1> (lbm/force.h:160-181) through a
    CoupledTaskSelector: each matching DYN_BGK_SCALAR_FORCE cell of `flow`
    takes SCALARFORCE += gbeta (scalar - rho_init) on the last axis.

    The buoyancy is formed inside the flow block's collision from the scalar
    field the scalar block's last step wrote, which is what FreeLB's loop
    (clear SCALARFORCE, coupling pass, flow step, scalar step;
    examples/cazsblock2d/cazsblock2d.cpp:385-409) computes, without the two
    extra passes over the domain.
    """

    def __init__(self, flow: BlockLattice, scalar: BlockLattice):
        self.flow, self.scalar = flow, scalar

    def getLat0(self) -> BlockLattice:
        return self.flow

    def getLat1(self) -> BlockLattice:
        return self.scalar

    def add_buoyancy(self, mask: int, gbeta: float, rho_init: float) -> None:
        """Key_TypePair<mask, force::Buoyancy<NSCELL, CELL1>>; GBETA / RHOINIT values."""
        abi.check(abi.lib().flb_add_buoyancy(self.flow.handle, self.scalar.handle, int(mask) & 0xFF,
                                             float(gbeta), float(rho_init)))
        refs = getattr(self.flow, "_buoy_src", [])
        refs.append(self.scalar)
        self.flow._buoy_src = refs

    def clear(self) -> None:
        abi.check(abi.lib().flb_clear_buoyancy(self.flow.handle))
        self.flow._buoy_src = []

    def step(self, n: int = 1) -> None:
        """n iterations of the coupled loop: the flow step, then the scalar step."""
        for _ in range(int(n)):
            self.flow.step(1)
            self.scalar.step(1)


class BlockHalo:
    """Same-level communication between the blocks of one domain (FreeLB's
    normalCommunicate / normalFullCommunicate, data_struct/block_lattice.hh:1669-1703).
    origins[k] is the global index of block k's local cell (0, 0, 0); a
    periodic domain axis of `extent` cells wraps from block to block.
    exchange() follows every step of all blocks."""

    def __init__(self, blocks: Sequence[BlockLattice], origins, periodic=(False, False, False),
                 extent=(0, 0, 0), stream=None):
        L = abi.lib()
        self.blocks = list(blocks)
        arr = (C.c_void_p * len(self.blocks))(*[b.handle for b in self.blocks])
        org = np.ascontiguousarray(np.asarray(origins, dtype=np.int32).reshape(-1, 3))
        assert org.shape[0] == len(self.blocks)
        per = np.ascontiguousarray(np.asarray(periodic, dtype=np.int32).reshape(3))
        ext = np.ascontiguousarray(np.asarray(extent, dtype=np.int32).reshape(3))
        h = C.c_void_p()
        abi.check(L.flb_halo_create(arr, len(self.blocks), abi.ptr(org, C.c_int), abi.ptr(per, C.c_int),
                                    abi.ptr(ext, C.c_int), stream, C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def info(self):
        n, cells = C.c_int(), C.c_int64()
        abi.check(abi.lib().flb_halo_info(self._h, C.byref(n), C.byref(cells)))
        return {"items": n.value, "ghost_cells": cells.value}

    def exchange(self) -> None:
        abi.check(abi.lib().flb_halo_exchange(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None):
            abi.lib().flb_halo_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_vtm(blocks, min_centres, voxels, directory: str, name: str, step: int = -1, strips=None,
              rho_name: str | None = "Rho", u_name: str | None = "Velocity", flag_name: str | None = None) -> None:
    """FreeLB's vtmo::vtmWriter::WriteBinary(step) for these blocks: one .vti
    per block (its RHO / VELOCITY fields and flags) and a .vtm listing them.
    min_centres[k] is the centre of block k's cell (0, 0, 0), voxels[k] its
    cell size, strips[k] the layers cut off each side."""
    blocks = list(blocks)
    n = len(blocks)
    arr = (C.c_void_p * n)(*[b.handle for b in blocks])
    mc = np.ascontiguousarray(np.asarray(min_centres, dtype=np.float64).reshape(n, 3))
    vx = np.ascontiguousarray(np.asarray(voxels, dtype=np.float64).reshape(n))
    st = None if strips is None else np.ascontiguousarray(np.asarray(strips, dtype=np.int32).reshape(n))
    enc = (lambda v: None if v is None else v.encode())
    abi.check(abi.lib().flb_write_vtm(arr, n, abi.ptr(mc, C.c_double), abi.ptr(vx, C.c_double),
                                      None if st is None else abi.ptr(st, C.c_int), directory.encode(),
                                      name.encode(), int(step), enc(rho_name), enc(u_name), enc(flag_name)))


class Refinement:
    """A coarse block and a fine block at twice its resolution whose interior
    covers the coarse cells [origin, origin + m) per axis (fine extent 2m + 4,
    overlap 1), exchanged as FreeLB's FullCommunicate(count) does for a
    refined pair (averCommunicate / intpCommunicate,
    data_struct/block_lattice.hh:1724-1811). run(n) is n counts of
    examples/cavref2d/cavref2d.cpp's loop: fine step, coarse step on even
    counts, communicate(count)."""

    def __init__(self, coarse: BlockLattice, fine: BlockLattice, origin, stream=None):
        L = abi.lib()
        self.coarse, self.fine = coarse, fine
        o = (C.c_int * 3)(*[int(v) for v in (list(origin) + [0, 0])[:3]])
        h = C.c_void_p()
        abi.check(L.flb_refine_create(coarse.handle, fine.handle, o, stream, C.byref(h)))
        self._h = h
        self.count = 0

    def info(self):
        a, b = C.c_int64(), C.c_int64()
        abi.check(abi.lib().flb_refine_info(self._h, C.byref(a), C.byref(b)))
        return {"average": a.value, "interp": b.value}

    def set_levels(self, coarse_level: int, max_level: int) -> None:
        """Levels of the pair in a deeper hierarchy (default 0 / 1)."""
        abi.check(abi.lib().flb_refine_set_levels(self._h, int(coarse_level), int(max_level)))

    def average(self, count: int) -> None:
        abi.check(abi.lib().flb_refine_average(self._h, int(count)))

    def interpolate(self, count: int) -> None:
        abi.check(abi.lib().flb_refine_interpolate(self._h, int(count)))

    def communicate(self, count: int) -> None:
        abi.check(abi.lib().flb_refine_communicate(self._h, int(count)))

    def run(self, n: int) -> None:
        abi.check(abi.lib().flb_refine_run(self._h, int(self.count), int(n)))
        self.count += int(n)

    def close(self) -> None:
        if getattr(self, "_h", None):
            abi.lib().flb_refine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class RefinementManager:
    """Blocks of several refinement levels and the refined pairs between them
    (any number of patches and levels), stepped by count as FreeLB's
    BlockLatticeManager does (data_struct/block_lattice.hh:736-764,
    1462-1561): at count c each block of level L steps when
    c % 2^(Lmax - L) == 0 (ApplyCellDynamics(count), Stream(count)); then
    every pair averages into its coarse block (AverageCommunicate(count),
    gated by the coarse level); then every pair interpolates into its fine
    block (InterpolateCommunicate(count), gated by the fine level).
    levels[k] is the level of blocks[k]; a pair's fine block must be one
    level above its coarse block."""

    def __init__(self, blocks: Sequence[BlockLattice], levels: Sequence[int], refinements: Sequence[Refinement]):
        self.blocks, self.levels, self.refinements = list(blocks), [int(v) for v in levels], list(refinements)
        if len(self.levels) != len(self.blocks):
            raise ValueError("one level per block")
        self.max_level = max(self.levels)
        lev = {id(b): lv for b, lv in zip(self.blocks, self.levels)}
        for b, lv in zip(self.blocks, self.levels):
            b.set_level(lv, self.max_level)
        for r in self.refinements:
            lc, lf = lev.get(id(r.coarse)), lev.get(id(r.fine))
            if lc is None or lf is None or lf != lc + 1:
                raise ValueError("a refinement's blocks must be managed blocks one level apart")
            r.set_levels(lc, self.max_level)
        self.count = 0

    def run(self, n: int) -> None:
        for _ in range(int(n)):
            c = self.count
            for b in self.blocks:
                b.ApplyCellDynamics(c)
                b.Stream(c)
            for r in self.refinements:
                r.average(c)
            for r in self.refinements:
                r.interpolate(c)
            self.count += 1


class BlockLatticeManager:
    """Many BlockLattice blocks stepped together, one launch per time step
    (FreeLB's BlockLatticeManager::ApplyCellDynamics / Stream over its blocks,
    data_struct/block_lattice.hh:744-778). The blocks keep their own set-up
    and read-out and must stay open while the manager lives."""

    def __init__(self, blocks: Sequence[BlockLattice], stream=None, halo: "BlockHalo" = None):
        L = abi.lib()
        self.blocks = list(blocks)
        arr = (C.c_void_p * len(self.blocks))(*[b.handle for b in self.blocks])
        h = C.c_void_p()
        abi.check(L.flb_batch_create(arr, len(self.blocks), stream, C.byref(h)))
        self._h = h
        self.halo = halo
        if halo is not None:
            abi.check(L.flb_batch_set_halo(self._h, halo.handle))

    def close(self) -> None:
        if getattr(self, "_h", None):
            abi.lib().flb_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def info(self):
        n, cells, cfg = C.c_int(), C.c_int64(), abi.LaunchCfg()
        abi.check(abi.lib().flb_batch_info(self._h, C.byref(n), C.byref(cells), C.byref(cfg)))
        return {"blocks": n.value, "cells": cells.value,
                "waves_per_group": cfg.waves_per_group, "grid": cfg.grid}

    def set_launch(self, waves_per_group: int, grid: int = 0) -> None:
        cfg = abi.LaunchCfg(int(waves_per_group), int(grid))
        abi.check(abi.lib().flb_batch_set_launch(self._h, C.byref(cfg)))

    def step(self, n: int = 1) -> None:
        abi.check(abi.lib().flb_batch_step(self._h, int(n)))

    def capture(self, n: int) -> None:
        abi.check(abi.lib().flb_batch_capture(self._h, int(n)))

    def replay(self) -> None:
        abi.check(abi.lib().flb_batch_replay(self._h))

    def synchronize(self) -> None:
        abi.check(abi.lib().flb_batch_synchronize(self._h))

    def autotune(self, steps: int = 10):
        cfg, ms = abi.LaunchCfg(), C.c_double()
        abi.check(abi.lib().flb_batch_autotune(self._h, int(steps), C.byref(cfg), C.byref(ms)))
        return {"waves_per_group": cfg.waves_per_group, "grid": cfg.grid, "ms_per_step": ms.value}
