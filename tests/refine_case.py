"""Helpers for the two-level refined cavity: the FreeLB fixture reader, the
mapping of FreeLB's blocks onto the oracle's (one coarse block over the whole
cavity, one fine block), and the same case handed to the GPU path."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O


def load_fixture(path):
    z = np.load(path)
    n, cell, p0, p1, counts = (int(v) for v in z["case"])
    blocks = []
    for k in range(int(z["nblocks"])):
        blocks.append({key: z[f"b{k}_{key}"] for key in ("level", "overlap", "dims", "omega", "min", "flags",
                                                          "pops", "fields")})
    return dict(n=n, cell=cell, p0=p0, p1=p1, counts=counts, blocks=blocks)


def load_fixture3(path):
    """A three-level FreeLB refined run and the oracle hierarchy of the same
    case: level-1 patch over block cells [p0, p1), level-2 patch over
    [q0, q1) (both per axis, `cell` coarse cells per block cell)."""
    z = np.load(path)
    n, cell, p0, p1, counts, q0, q1 = (int(v) for v in z["case"])
    blocks = []
    for k in range(int(z["nblocks"])):
        blocks.append({key: z[f"b{k}_{key}"] for key in ("level", "overlap", "dims", "omega", "min", "flags",
                                                          "pops", "fields")})
    o1, m1 = 1 + p0 * cell, (p1 - p0) * cell
    o2, m2 = 2 + 2 * ((1 + q0 * cell) - o1), 2 * (q1 - q0) * cell
    patches = [(0, (o1, o1), m1), (1, (o2, o2), m2)]
    return dict(n=n, counts=counts, patches=patches, blocks=blocks)


def fixture_mismatches(rc, fx, fine_mask=None):
    """Cells where the oracle's state differs from FreeLB's: every non-void
    cell of each level-0 block (base and overlap, which mirror the coarse
    block's cells, the patch's outer coarse layer included) and every cell of
    the level-1 block (or those of fine_mask), populations and RHO / VELOCITY
    fields."""
    d, q = rc.d, rc.q
    cd = rc.cdims
    popsC = rc.popsC.reshape((q,) + cd[::-1])
    fldC = rc.coarse.fields().astype(np.float32).reshape((d + 1,) + cd[::-1])
    bad, checked = 0, 0
    for b in fx["blocks"]:
        nx, ny, nz = (int(v) for v in b["dims"])
        P = b["pops"].reshape((q, nz, ny, nx))
        Fd = b["fields"].reshape((d + 1, nz, ny, nx))
        fl = b["flags"].reshape((nz, ny, nx))
        if int(b["level"]) == 0:
            g0 = [int(round(float(b["min"][a]))) + 1 for a in range(d)] + [0] * (3 - d)
            live = (fl & O.VOID) == 0
            sl = (slice(g0[2], g0[2] + nz), slice(g0[1], g0[1] + ny), slice(g0[0], g0[0] + nx))
            mine = popsC[(slice(None),) + sl]
            mf = fldC[(slice(None),) + sl]
            diff = np.any(mine != P, axis=0) | (np.any(mf != Fd, axis=0) & (fl == O.AABB))
            bad += int((diff & live).sum())
            checked += int(live.sum())
        else:
            assert tuple(int(v) for v in b["dims"]) == rc.fdims
            assert np.float32(b["omega"]) == np.float32(rc.omega_f)
            mine = rc.popsF.reshape(P.shape)
            mf = rc.fine.fields().astype(np.float32).reshape(Fd.shape)
            sel = np.ones((nz, ny, nx), dtype=bool) if fine_mask is None else fine_mask.reshape((nz, ny, nx))
            bad += int(((np.any(mine != P, axis=0) | np.any(mf != Fd, axis=0)) & sel).sum())
            checked += int(sel.sum())
    return bad, checked


def gpu_refined(rc, storage="f32"):
    """The oracle's coarse and fine lattices as GPU blocks plus their
    Refinement, at the oracle's current state. The fine block's outer ghost
    layer is its overlap (not updated); its inner ghost layer is updated."""
    from freelb_amd import BlockLattice, Refinement
    from gpu_case import gpu_lattice
    coarse = gpu_lattice(rc.coarse, storage, rc.popsC)
    fl = rc.fine.flags.reshape(rc.fdims[::-1]).copy()
    if rc.d == 3:
        fl[0, :, :] = fl[-1, :, :] = O.COVERED
        fl[:, 0, :] = fl[:, -1, :] = O.COVERED
        fl[:, :, 0] = fl[:, :, -1] = O.COVERED
    else:
        fl[0, 0, :] = fl[0, -1, :] = O.COVERED
        fl[0, :, 0] = fl[0, :, -1] = O.COVERED
    fine_case = O.Lattice(rc.q, rc.fdims, fl.reshape(-1), rc.fine.tasks, [], omega=rc.fine.omega,
                          rho_field=rc.fine.buffers["rho_field"], u_field=rc.fine.buffers["u_field"])
    fine = gpu_lattice(fine_case, storage, rc.popsF)
    ref = Refinement(coarse, fine, rc.origin)
    return coarse, fine, ref


def compare_masks(rc):
    """Cells whose state the GPU must reproduce: coarse cells that are updated
    and not covered (live cells and the patch's outer coarse layer), and the
    fine interior with its inner ghost layer."""
    cm = (rc.coarse.flags != O.VOID) & (rc.coarse.flags != O.COVERED)
    fm = np.zeros(rc.fdims[::-1], dtype=bool)
    if rc.d == 3:
        fm[1:-1, 1:-1, 1:-1] = True
    else:
        fm[0, 1:-1, 1:-1] = True
    return cm, fm.reshape(-1)


def gpu_hierarchy(rh, storage="f32"):
    """The oracle hierarchy's blocks as GPU blocks (each patch block's outer
    ghost layer is its overlap), its pairs as Refinements and the manager
    that steps them, at the oracle's current state."""
    from freelb_amd import Refinement, RefinementManager
    from gpu_case import gpu_lattice
    blocks = [gpu_lattice(rh.blocks[0], storage, rh.pops[0])]
    for k in range(1, len(rh.blocks)):
        b = rh.blocks[k]
        fl = b.flags.reshape(rh.dims[k][::-1]).copy()
        if rh.d == 3:
            fl[0, :, :] = fl[-1, :, :] = O.COVERED
            fl[:, 0, :] = fl[:, -1, :] = O.COVERED
            fl[:, :, 0] = fl[:, :, -1] = O.COVERED
        else:
            fl[0, 0, :] = fl[0, -1, :] = O.COVERED
            fl[0, :, 0] = fl[0, :, -1] = O.COVERED
        case = O.Lattice(rh.q, rh.dims[k], fl.reshape(-1), b.tasks, [], omega=b.omega,
                         rho_field=b.buffers["rho_field"], u_field=b.buffers["u_field"])
        blocks.append(gpu_lattice(case, storage, rh.pops[k]))
    refs = [Refinement(blocks[pa], blocks[ch], o) for pa, ch, o in rh.patches]
    return blocks, refs, RefinementManager(blocks, rh.levels, refs)


def hierarchy_masks(rh):
    """Cells whose state the GPU must reproduce: per block, the updated cells
    (the cavity without its void shell, a patch without its outer ghost
    layer) that no finer patch covers."""
    out = []
    for k, b in enumerate(rh.blocks):
        m = (b.flags != O.VOID) & (b.flags != O.COVERED)
        if k > 0:
            inner = np.zeros(rh.dims[k][::-1], dtype=bool)
            if rh.d == 3:
                inner[1:-1, 1:-1, 1:-1] = True
            else:
                inner[0, 1:-1, 1:-1] = True
            m &= inner.reshape(-1)
        out.append(m)
    return out


def hierarchy_fixture_mismatches(rh, fx):
    """Cells where the oracle hierarchy's state differs from FreeLB's refined
    run with several levels. Each FreeLB block of level L is placed on the
    hierarchy's level-L block by its min corner (FreeLB's level-L blocks
    tile that block). Compared: every non-void cell of a FreeLB block that no
    other FreeLB block of its level holds as a base cell (its own base
    cells, and overlap cells facing a coarser level or the domain edge);
    overlap cells that mirror another block's base cells are compared in
    that block. Of an overlap cell that no block owns (it faces a coarser or
    finer level), only the populations that stream into the block's base
    cells are compared: FreeLB's communication fills those and leaves the
    others (which stream into overlap cells) stale, and two blocks holding
    such a cell then disagree on them. Cells the hierarchy marks COVERED by a
    finer patch are not compared (FreeLB's overlap cells there hold
    averages). Populations, and the RHO / VELOCITY fields on fluid cells."""
    d, q = rh.d, rh.q
    mins, vox = [np.array([-1.0] * d)], [1.0]
    for pa, ch, o in rh.patches:
        v = vox[pa] / 2
        mins.append(mins[pa] + vox[pa] * np.array(o[:d], dtype=float) - 2 * v)
        vox.append(v)

    def place(b):
        L = int(b["level"])
        nx, ny, nz = (int(v) for v in b["dims"])
        for k in range(len(rh.blocks)):
            if rh.levels[k] != L:
                continue
            off = np.rint((np.asarray(b["min"][:d], dtype=float) - mins[k]) / vox[k]).astype(int)
            if all(0 <= off[a] and off[a] + (nx, ny, nz)[a] <= rh.dims[k][a] for a in range(d)):
                z0 = int(off[2]) if d == 3 else 0
                return k, (slice(z0, z0 + nz), slice(int(off[1]), int(off[1]) + ny),
                           slice(int(off[0]), int(off[0]) + nx))
        raise AssertionError(f"FreeLB block of level {L} at {b['min']} has no place in the hierarchy")

    placed = [place(b) for b in fx["blocks"]]
    # base cells of every FreeLB block, per hierarchy block
    owner = [np.full(rh.dims[k][::-1], -1, dtype=int) for k in range(len(rh.blocks))]
    for j, (b, (k, sl)) in enumerate(zip(fx["blocks"], placed)):
        ov = int(b["overlap"])
        base = np.zeros(tuple(s.stop - s.start for s in sl), dtype=bool)
        inner = tuple(slice(ov, n - ov) if n > 1 else slice(None) for n in base.shape)
        base[inner] = True
        owner[k][sl][base] = j
    c = _velocities(q)
    bad = checked = 0
    for j, (b, (k, sl)) in enumerate(zip(fx["blocks"], placed)):
        nx, ny, nz = (int(v) for v in b["dims"])
        dims = rh.dims[k]
        ov = int(b["overlap"])
        P = b["pops"].reshape((q, nz, ny, nx))
        Fd = b["fields"].reshape((d + 1, nz, ny, nx))
        fl = b["flags"].reshape((nz, ny, nx))
        mine = rh.pops[k].reshape((q,) + dims[::-1])[(slice(None),) + sl]
        mf = rh.blocks[k].fields().astype(np.float32).reshape((d + 1,) + dims[::-1])[(slice(None),) + sl]
        mfl = rh.blocks[k].flags.reshape(dims[::-1])[sl]
        own = owner[k][sl]
        live = ((fl & O.VOID) == 0) & (mfl != O.COVERED)
        base = np.zeros((nz, ny, nx), dtype=bool)
        base[tuple(slice(ov, n - ov) if n > 1 else slice(None) for n in (nz, ny, nx))] = True
        # per direction: the populations compared in each cell
        zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        for i in range(q):
            tx, ty, tz = xx + c[i][0], yy + c[i][1], zz + (c[i][2] if d == 3 else 0)
            inside = (tx >= 0) & (tx < nx) & (ty >= 0) & (ty < ny) & (tz >= 0) & (tz < nz)
            feeds = np.zeros_like(base)
            feeds[inside] = base[tz[inside], ty[inside], tx[inside]]
            use = live & ((own == j) | ((own < 0) & (base | feeds)))
            bad += int(((mine[i] != P[i]) & use).sum())
            checked += int(use.sum())
        fuse = live & (own == j) & (fl == O.AABB)
        bad += int((np.any(mf != Fd, axis=0) & fuse).sum())
    return bad, checked


def _velocities(q):
    """FreeLB's D2Q9 / D3Q19 velocity order (lbm/lattice_set.h:168-194, 259-296)."""
    if q == 9:
        return [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (1, 1, 0), (-1, -1, 0), (1, -1, 0),
                (-1, 1, 0)]
    return [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0),
            (-1, -1, 0), (1, 0, 1), (-1, 0, -1), (0, 1, 1), (0, -1, -1), (1, -1, 0), (-1, 1, 0), (1, 0, -1),
            (-1, 0, 1), (0, 1, -1), (0, -1, 1)]
