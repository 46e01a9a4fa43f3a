"""Block decomposition of a domain box, as FreeLB's BlockGeometry2D / 3D builds
its blocks (geometry/block_geometry3d.hh:156-171, CreateBlocks 489-510,
DivideBlocks 477-487), and the GPU blocks plus halo that carry it.

A domain of nx x ny x nz cells (one ghost layer around it, FreeLB's
overlap 1) is cut into blocks either by explicit counts per axis
(AABB::divide, geometry/basic_geometry.hh:192-250: each axis in equal parts,
the last `remainder` parts one cell longer) or by a total block count, for
which the counts per axis are chosen as DivideBlock3D / DivideBlock2D choose
them (geometry/basic_geometry.h:654-700 in 3D, the factorisation whose
block shape is closest to a cube; 489-514 in 2D, the count ratio closest to
the domain's aspect ratio). FreeLB's split of a count that no such
factorisation fills exactly (its "rest" branches) is not reproduced:
such counts raise ValueError.

Each block is one BlockLattice over its cells plus one ghost layer; the
ghost layers mirror the neighbouring blocks' cells through a BlockHalo
(FreeLB's InitComm / normalCommunicate, block_geometry3d.hh:531-605,
data_struct/block_lattice.hh:1669-1703). Flags, fields and populations are
given on the whole domain (ghost layer included) and cut per block.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np

from . import abi
from .lattice import BlockHalo, BlockLattice, HalfwayBB, Key_TypePair


def divide_axis(n: int, parts: int):
    """[lo, hi) intervals of AABB::divide along one axis of n cells."""
    if parts < 1 or parts > n:
        raise ValueError(f"cannot divide {n} cells into {parts} parts")
    base, rem = divmod(n, parts)
    out, lo = [], 0
    for k in range(parts):
        size = base + (1 if k >= parts - rem else 0)
        out.append((lo, lo + size))
        lo += size
    return out


def _ratio(nx, ny, nz, ix, iy, iz):
    """DivideBlock3D's aspect measure in float (an axis cut into more parts
    than cells gives inf, as there)."""
    f = np.float32
    a, b, c = f(nx // ix), f(ny // iy), f(nz // iz)
    with np.errstate(divide="ignore", invalid="ignore"):
        return (a / b - f(1)) ** 2 + (b / c - f(1)) ** 2 + (c / a - f(1)) ** 2


def block_counts(dims: Sequence[int], blocknum: int):
    """Blocks per axis for `blocknum` blocks (DivideBlock3D's choice; 2D
    domains, nz == 1, divide x and y only)."""
    nx, ny, nz = (int(v) for v in dims)
    if blocknum == 1:
        return (1, 1, 1)
    if nz == 1:
        # DivideBlock2D (geometry/basic_geometry.h:489-514): the x / y count
        # ratio closest to nx / ny, later candidates winning ties
        f = np.float32
        target = f(nx) / f(ny)
        dif = np.abs(target - f(1)) + f(1)
        bx = by = 0
        for i in range(1, blocknum + 1):
            j = blocknum // i
            dist = np.abs(target - f(i) / f(j))
            if dist <= dif:
                dif, bx, by = dist, i, j
        counts = (bx, by, 1)
    else:
        bx, by, bz = 1, 1, blocknum
        best = _ratio(nx, ny, nz, 1, 1, blocknum)
        for ix in range(1, blocknum + 1):
            iy = 1
            while iy * ix <= blocknum:
                iz = blocknum // (ix * iy)
                while iz * iy * ix <= blocknum:
                    if (ix + 1) * iy * iz > blocknum and ix * (iy + 1) * iz > blocknum:
                        r = _ratio(nx, ny, nz, ix, iy, iz)
                        if r < best:
                            best, bx, by, bz = r, ix, iy, iz
                    iz += 1
                iy += 1
        counts = (bx, by, bz)
    if int(np.prod(counts)) != blocknum:
        raise ValueError(f"{blocknum} blocks do not factor into a division of {dims} (counts {counts}); "
                         "give the counts per axis")
    return counts


class BlockGeometry:
    """The blocks of a domain box. dims: cells per axis (nz = 1 for 2D);
    blocknum or counts: the division; periodic: axes that wrap.

    Arrays of the whole domain are "padded": a non-periodic axis of n cells
    carries the domain's ghost layer (indices 0 and n + 1, cells 1 .. n), a
    periodic axis has none (cells 0 .. n - 1, wrapping). boxes[k] is block
    k's [lo, hi) per axis in those indices, its own ghost layer included;
    on a periodic axis lo may be -1 and hi n + 1 (wrapped). Blocks are
    numbered x fastest, then y, then z, as FreeLB numbers them."""

    def __init__(self, dims: Sequence[int], blocknum: int | None = None, counts: Sequence[int] | None = None,
                 periodic=(False, False, False)):
        self.dims = tuple(int(v) for v in dims)
        self.d = 2 if self.dims[2] == 1 else 3
        if counts is None:
            counts = block_counts(self.dims, 1 if blocknum is None else int(blocknum))
        self.counts = tuple(int(v) for v in counts)
        if self.d == 2 and self.counts[2] != 1:
            raise ValueError("a 2D domain has one block along z")
        self.periodic = tuple(bool(v) for v in periodic)
        if self.d == 2 and self.periodic[2]:
            raise ValueError("a 2D domain has no z axis to wrap")
        ivs = [divide_axis(n, c) for n, c in zip(self.dims, self.counts)]
        # per axis: first padded index of the block's first cell
        shift = [0 if (self.periodic[a] or (a == 2 and self.d == 2)) else 1 for a in range(3)]
        self.boxes = []
        for (z0, z1) in ivs[2]:
            for (y0, y1) in ivs[1]:
                for (x0, x1) in ivs[0]:
                    lo, hi = [], []
                    for a, (c0, c1) in enumerate(((x0, x1), (y0, y1), (z0, z1))):
                        if a == 2 and self.d == 2:
                            lo.append(0)
                            hi.append(1)
                        else:
                            lo.append(c0 + shift[a] - 1)
                            hi.append(c1 + shift[a] + 1)
                    self.boxes.append((tuple(lo), tuple(hi)))
        self.padded = tuple(1 if (a == 2 and self.d == 2) else (n if self.periodic[a] else n + 2)
                            for a, n in enumerate(self.dims))

    def __len__(self):
        return len(self.boxes)

    def block_dims(self, k):
        lo, hi = self.boxes[k]
        return tuple(b - a for a, b in zip(lo, hi))

    def _index(self, k, a):
        lo, hi = self.boxes[k]
        idx = np.arange(lo[a], hi[a])
        return idx % self.padded[a] if self.periodic[a] else idx

    def cut(self, arr, k, comps=1):
        """Block k's part of a padded-domain array [comps][N] (x fastest)."""
        nx, ny, nz = self.padded
        a = np.asarray(arr).reshape(comps, nz, ny, nx)
        ix, iy, iz = (self._index(k, ax) for ax in range(3))
        return np.ascontiguousarray(a[:, iz][:, :, iy][:, :, :, ix].reshape(comps, -1))

    def paste(self, parts, comps=1, dtype=np.float32):
        """The padded-domain array [comps][N] assembled from each block's
        cells (its ghost layer dropped); the domain's ghost layer stays 0."""
        nx, ny, nz = self.padded
        out = np.zeros((comps, nz, ny, nx), dtype=dtype)
        for k, p in enumerate(parts):
            bx, by, bz = self.block_dims(k)
            v = np.asarray(p).reshape(comps, bz, by, bx)
            sl = [slice(1, -1), slice(1, -1), slice(1, -1) if self.d == 3 else slice(0, 1)]
            ix, iy, iz = (self._index(k, ax)[sl[ax]] for ax in range(3))
            out[np.ix_(np.arange(comps), iz, iy, ix)] = v[:, sl[2], sl[1], sl[0]]
        return out.reshape(comps, -1)

    def build(self, q: int, flags, tasks, bcs=(), omega: float = 1.0, storage: str = "f32", rho_field=None,
              u_field=None, pops=None, mrt_rates=None, stream=None):
        """The blocks as GPU lattices and their halo. flags (and the optional
        RHO / VELOCITY fields and populations) cover the padded domain;
        tasks are Key_TypePair or (mask, dynamics[, C_s]) tuples, bcs
        HalfwayBB or (cell mask, void mask, kind) tuples. Returns
        (blocks, halo); closing both is the caller's."""
        d = 3 if q in (19, 7) else 2
        tk = [t if isinstance(t, Key_TypePair) else Key_TypePair(t[0], t[1], float(t[2]) if len(t) > 2 else 0.0)
              for t in tasks]
        bc = [b if isinstance(b, HalfwayBB) else HalfwayBB(b[0], b[1], b[2] == abi.HWBC_MOVING,
                                                          b[2] == abi.HWBC_ANTI) for b in bcs]
        blocks = []
        try:
            for k in range(len(self.boxes)):
                lat = BlockLattice(q, self.block_dims(k), omega=omega, storage=storage, stream=stream)
                blocks.append(lat)
                lat.set_ghost_overlap(True)
                lat.upload_flags(self.cut(flags, k)[0].astype(np.uint8))
                lat.set_tasks(tk)
                if bc:
                    lat.set_halfway_bc(bc)
                if rho_field is not None:
                    lat.set_field_rho(self.cut(rho_field, k)[0])
                if u_field is not None:
                    lat.set_field_u(self.cut(u_field, k, d).reshape(-1))
                if mrt_rates is not None:
                    lat.set_mrt_rates(mrt_rates)
                if pops is not None:
                    lat.upload_pops(self.cut(pops, k, q))
            origins = [lo for lo, _ in self.boxes]
            extent = tuple(n if p else 0 for n, p in zip(self.padded, self.periodic))
            halo = BlockHalo(blocks, origins, periodic=self.periodic, extent=extent, stream=stream)
        except Exception:
            for b in blocks:
                b.close()
            raise
        return blocks, halo
