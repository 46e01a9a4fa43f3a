"""ctypes binding of the C oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module; the product (freelb_amd/) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

(DYN_NONE, DYN_BGK, DYN_BGK_FIELD, DYN_BOUNCEBACK, DYN_BOUNCEBACK_MW, DYN_MRT, DYN_SMAG_BGK, DYN_SMAG_MRT,
 DYN_BGK_FORCE, DYN_BGK_WRITE, DYN_BGK_SCALAR_FORCE, DYN_SCALAR_BGK, DYN_SCALAR_BGK_FIELD) = range(13)
HWBC_NORMAL, HWBC_MOVING, HWBC_ANTI = 0, 1, 2

# Cavity flags of the reference drivers (benchmarks/cavity3d/cavity3d.cpp:121-124).
VOID, AABB, BB, BBMW = 1, 2, 4, 8
# Coarse cells covered by a finer patch beyond its outer coarse layer (no task).
COVERED = 16
# Body-force density of the 'force' cavity (float values, as FreeLB's CONSTFORCE<float>);
# the reference harness takes the same decimal strings.
FORCE_STR = ("1e-4", "-5e-5", "2e-5")
FORCE = tuple(np.float32(v) for v in FORCE_STR)


class _Task(C.Structure):
    _fields_ = [("mask", C.c_uint8), ("dynamics", C.c_uint8), ("pad", C.c_uint8 * 2),
                ("smagorinsky", C.c_double)]


class _HWBC(C.Structure):
    _fields_ = [("cell_mask", C.c_uint8), ("void_mask", C.c_uint8), ("kind", C.c_uint8),
                ("pad", C.c_uint8)]


class _Lattice(C.Structure):
    _fields_ = [("q", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("overlap", C.c_int), ("flags", C.POINTER(C.c_uint8)), ("ntasks", C.c_int),
                ("tasks", _Task * 8), ("nbc", C.c_int), ("bcs", _HWBC * 4),
                ("omega", C.c_double), ("rho_field", C.POINTER(C.c_double)),
                ("u_field", C.POINTER(C.c_double)), ("mrt_rates", C.POINTER(C.c_double)),
                ("nthreads", C.c_int), ("force", C.c_double * 3),
                ("sforce", C.POINTER(C.c_double)), ("u_ext", C.POINTER(C.c_double))]


class _Buoyancy(C.Structure):
    _fields_ = [("scalar_rho", C.POINTER(C.c_double)), ("mask", C.c_uint8), ("pad", C.c_uint8 * 7),
                ("gbeta", C.c_double), ("rho_init", C.c_double)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        for name, T in (("f32", C.c_float), ("f64", C.c_double)):
            P = C.POINTER(T)
            getattr(_lib, "flbo_run_" + name).argtypes = [C.POINTER(_Lattice), P, P, C.c_int]
            getattr(_lib, "flbo_run_" + name).restype = C.c_int
            getattr(_lib, "flbo_rhou_" + name).argtypes = [C.POINTER(_Lattice), P, C.c_uint8, P]
            getattr(_lib, "flbo_init_feq_" + name).argtypes = [C.POINTER(_Lattice), P, P]
            P2 = C.POINTER(C.c_int)
            getattr(_lib, "flbo_refine_average_" + name).argtypes = [C.POINTER(_Lattice), P, C.POINTER(_Lattice),
                                                                     P, P2]
            getattr(_lib, "flbo_refine_interpolate_" + name).argtypes = [C.POINTER(_Lattice), P,
                                                                         C.POINTER(_Lattice), P, P2]
        for name, T in (("f32", C.c_float), ("f64", C.c_double)):
            P = C.POINTER(T)
            PP = C.POINTER(P)
            fn = getattr(_lib, "flbo_run_coupled_" + name)
            fn.argtypes = [C.POINTER(_Lattice), P, P, C.POINTER(C.POINTER(_Lattice)), PP, PP,
                           C.POINTER(_Buoyancy), C.c_int, C.c_int, C.c_int]
            fn.restype = C.c_int
        _lib.flbo_cavity_flags.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint8)]
        _lib.flbo_convert_from_rt.argtypes = [C.c_float, C.c_float, C.c_float,
                                              C.POINTER(C.c_float), C.POINTER(C.c_float)]
    return _lib


def _ptr(a, T):
    return a.ctypes.data_as(C.POINTER(T))


class Lattice:
    """One block of the reference lattice: flags, task table, halfway BCs, fields."""

    def __init__(self, q, dims, flags, tasks, bcs=(), omega=1.0, rho_field=None, u_field=None,
                 mrt_rates=None, overlap=1, nthreads=0, force=(0.0, 0.0, 0.0)):
        self.q = q
        self.nx, self.ny, self.nz = dims
        self.N = self.nx * self.ny * self.nz
        self.flags = np.ascontiguousarray(flags, dtype=np.uint8).reshape(-1)
        assert self.flags.size == self.N
        # python-side copies, so a test can hand the same case to the GPU path
        self.tasks = [tuple(t) for t in tasks]
        self.bcs = [tuple(b) for b in bcs]
        self.omega = float(omega)
        self.rho_field = None if rho_field is None else np.asarray(rho_field, dtype=np.float64)
        self.u_field = None if u_field is None else np.asarray(u_field, dtype=np.float64)
        self.mrt_rates = None if mrt_rates is None else np.asarray(mrt_rates, dtype=np.float64)
        self._keep = [self.flags]
        L = _Lattice()
        L.q, L.nx, L.ny, L.nz, L.overlap = q, self.nx, self.ny, self.nz, overlap
        L.flags = _ptr(self.flags, C.c_uint8)
        assert len(tasks) <= 8 and len(bcs) <= 4
        L.ntasks = len(tasks)
        for k, t in enumerate(tasks):
            mask, dyn = t[0], t[1]
            L.tasks[k].mask, L.tasks[k].dynamics = mask, dyn
            L.tasks[k].smagorinsky = t[2] if len(t) > 2 else 0.0
        L.nbc = len(bcs)
        for k, (cm, vm, kind) in enumerate(bcs):
            L.bcs[k].cell_mask, L.bcs[k].void_mask, L.bcs[k].kind = cm, vm, kind
        L.omega = float(omega)
        # the arrays the C side reads (and FLBO_DYN_BGK_WRITE cells write)
        self.buffers = {}
        for name, arr in (("rho_field", rho_field), ("u_field", u_field), ("mrt_rates", mrt_rates)):
            if arr is not None:
                a = np.array(arr, dtype=np.float64).reshape(-1)
                self._keep.append(a)
                self.buffers[name] = a
                setattr(L, name, _ptr(a, C.c_double))
        L.nthreads = nthreads
        self.force = tuple(float(v) for v in force) + (0.0,) * (3 - len(force))
        for a in range(3):
            L.force[a] = self.force[a]
        self.L = L

    @property
    def d(self):
        return 3 if self.q in (19, 7) else 2

    def run(self, pops, steps):
        """Advance pops[q][N] (float32 or float64, natural layout) in place."""
        T = C.c_float if pops.dtype == np.float32 else C.c_double
        name = "f32" if pops.dtype == np.float32 else "f64"
        assert pops.flags.c_contiguous and pops.size == self.q * self.N
        scratch = np.empty_like(pops)
        rc = getattr(lib(), "flbo_run_" + name)(C.byref(self.L), _ptr(pops, T), _ptr(scratch, T), steps)
        if rc != 0:
            raise RuntimeError(f"flbo_run_{name} failed ({rc})")
        return pops

    def rhou(self, pops, mask=0xFF & ~VOID):
        T = C.c_float if pops.dtype == np.float32 else C.c_double
        name = "f32" if pops.dtype == np.float32 else "f64"
        out = np.zeros((self.d + 1, self.N), dtype=pops.dtype)
        getattr(lib(), "flbo_rhou_" + name)(C.byref(self.L), _ptr(pops, T), mask, _ptr(out, T))
        return out

    def init_feq(self, rhou):
        rhou = np.ascontiguousarray(rhou)
        T = C.c_float if rhou.dtype == np.float32 else C.c_double
        name = "f32" if rhou.dtype == np.float32 else "f64"
        pops = np.empty((self.q, self.N), dtype=rhou.dtype)
        getattr(lib(), "flbo_init_feq_" + name)(C.byref(self.L), _ptr(rhou, T), _ptr(pops, T))
        return pops


    def fields(self):
        """Current RHO and VELOCITY fields, [(d+1)][N] (written by BGK_WRITE cells)."""
        rho = self.buffers.get("rho_field", np.ones(self.N))
        u = self.buffers.get("u_field", np.zeros(self.d * self.N)).reshape(self.d, self.N)
        return np.concatenate([rho.reshape(1, -1), u], axis=0)


def run_coupled(flow, pops, scalars, spops, couplings, steps):
    """The coupled loop (flbo_run_coupled): flow lattice with a SCALARFORCE
    field, scalar lattices advected by the flow's VELOCITY field, buoyancy
    couplings (scalar index, mask, gbeta, rho_init). pops / spops are advanced
    in place; the flow's fields and SCALARFORCE and the scalars' RHO fields
    are updated in their buffers."""
    T = C.c_float if pops.dtype == np.float32 else C.c_double
    name = "f32" if pops.dtype == np.float32 else "f64"
    if "u_field" not in flow.buffers or "rho_field" not in flow.buffers:
        raise ValueError("the flow lattice needs RHO and VELOCITY fields")
    sforce = flow.buffers.setdefault("sforce", np.zeros(flow.N))
    flow.L.sforce = _ptr(sforce, C.c_double)
    uf = flow.buffers["u_field"]
    for sc in scalars:
        sc.L.u_ext = _ptr(uf, C.c_double)
        sc._keep.append(uf)
    n = len(scalars)
    LP = (C.POINTER(_Lattice) * max(1, n))(*[C.pointer(sc.L) for sc in scalars])
    scratch = [np.empty_like(p) for p in spops]
    PA = (C.POINTER(T) * max(1, n))(*[_ptr(p, T) for p in spops])
    PB = (C.POINTER(T) * max(1, n))(*[_ptr(p, T) for p in scratch])
    B = (_Buoyancy * max(1, len(couplings)))()
    for k, (si, mask, gbeta, rho0) in enumerate(couplings):
        B[k].scalar_rho = _ptr(scalars[si].buffers["rho_field"], C.c_double)
        B[k].mask, B[k].gbeta, B[k].rho_init = mask, float(gbeta), float(rho0)
    fscratch = np.empty_like(pops)
    rc = getattr(lib(), "flbo_run_coupled_" + name)(C.byref(flow.L), _ptr(pops, T), _ptr(fscratch, T), LP, PA,
                                                    PB, B, len(couplings), n, int(steps))
    if rc != 0:
        raise RuntimeError(f"flbo_run_coupled_{name} failed ({rc})")
    return pops


def refine_average(coarse, popsC, fine, popsF, origin):
    """averCommunicate into the outer coarse layer of the patch (in place on popsC)."""
    T = C.c_float if popsC.dtype == np.float32 else C.c_double
    name = "f32" if popsC.dtype == np.float32 else "f64"
    o = (C.c_int * 3)(*[int(v) for v in origin])
    getattr(lib(), "flbo_refine_average_" + name)(C.byref(coarse.L), _ptr(popsC, T), C.byref(fine.L),
                                                  _ptr(popsF, T), o)


def refine_interpolate(coarse, popsC, fine, popsF, origin):
    """intpCommunicate into both fine ghost layers (in place on popsF)."""
    T = C.c_float if popsC.dtype == np.float32 else C.c_double
    name = "f32" if popsC.dtype == np.float32 else "f64"
    o = (C.c_int * 3)(*[int(v) for v in origin])
    getattr(lib(), "flbo_refine_interpolate_" + name)(C.byref(coarse.L), _ptr(popsC, T), C.byref(fine.L),
                                                      _ptr(popsF, T), o)


class RefinedCavity:
    """The two-level lid-driven cavity of examples/cavref2d/cavref2d.cpp with a
    centred patch: FreeLB's block cells of `cell` cells per axis whose index
    lies in [p0, p1) on every axis are refined once. Here the coarse level is
    one block over the whole cavity (the reference's coarse blocks tile it and
    mirror each other through their overlap cells, which one block reproduces)
    whose cells inside the patch beyond its outer layer are COVERED, and the
    fine level one block of 2m + 4 cells per axis (two fine ghost layers).
    Fluid cells collide with BGK<rhoU<CELL, true>>, wall cells with
    BGK<useFieldrhoU>, halfway normal / moving walls (cavref2d.cpp:155-179).
    The fine omega is RefineConverter::getOmegaF (unit_converter.h:627)."""

    def __init__(self, lattice, Ni, Nj, Nk, cell, p0, p1, RT=0.6, visc=0.01, u_phys=0.05, dtype=np.float32):
        q = 19 if lattice == "d3q19" else 9
        d = 3 if q == 19 else 2
        nk = Nk if d == 3 else 0
        self.q, self.d, self.dtype = q, d, dtype
        flags = cavity_flags(Ni, Nj, nk)
        self.cdims = (Ni + 2, Nj + 2, (Nk + 2) if d == 3 else 1)
        self.origin = [1 + p0 * cell] * d + [0] * (3 - d)
        self.m = (p1 - p0) * cell
        f3 = flags.reshape(self.cdims[::-1])
        o, m = self.origin[0], self.m
        inner = (slice(o + 1, o + m - 1),) * d
        if d == 3:
            f3[inner[::-1]] = COVERED
        else:
            f3[0][inner[::-1]] = COVERED
        flags = f3.reshape(-1)
        omega, ulat = convert_from_rt(np.float32(RT), np.float32(visc), np.float32(u_phys))
        omega_f = np.float32(1) / (np.float32(2) / omega - np.float32(0.5))
        if dtype == np.float64:
            omega, omega_f = np.float64(1.0 / RT), 1.0 / (2.0 * RT - 0.5)
        self.omega_c, self.omega_f, self.ulat = omega, omega_f, ulat
        N = flags.size
        rho = np.ones(N)
        u = np.zeros((d, N))
        u[0, flags == BBMW] = float(ulat)
        tasks = [(AABB, DYN_BGK_WRITE), (BB | BBMW, DYN_BGK_FIELD)]
        bcs = [(BB, VOID, HWBC_NORMAL), (BBMW, VOID, HWBC_MOVING)]
        self.coarse = Lattice(q, self.cdims, flags, tasks, bcs, omega=float(omega), rho_field=rho, u_field=u)
        fn = 2 * m + 4
        self.fdims = (fn, fn, fn if d == 3 else 1)
        NF = int(np.prod(self.fdims))
        self.fine = Lattice(q, self.fdims, np.full(NF, AABB, dtype=np.uint8), tasks, [], omega=float(omega_f),
                            rho_field=np.ones(NF), u_field=np.zeros((d, NF)))
        w = weights(q, dtype)
        self.popsC = np.ascontiguousarray(np.repeat(w[:, None], N, axis=1).astype(dtype))
        self.popsF = np.ascontiguousarray(np.repeat(w[:, None], NF, axis=1).astype(dtype))
        self.count = 0

    def run(self, n):
        """n counts: fine step, coarse step on even counts, FullCommunicate(count)."""
        for _ in range(n):
            c = self.count
            self.fine.run(self.popsF, 1)
            if c % 2 == 0:
                self.coarse.run(self.popsC, 1)
                refine_average(self.coarse, self.popsC, self.fine, self.popsF, self.origin)
            refine_interpolate(self.coarse, self.popsC, self.fine, self.popsF, self.origin)
            self.count += 1


class RefinedHierarchy:
    """A lid-driven cavity refined in several patches and levels, stepped as
    FreeLB's BlockLatticeManager loop does for any number of levels
    (data_struct/block_lattice.hh:736-764, 1462-1561): at count c every block
    of level L with c % 2^(Lmax - L) == 0 steps (ApplyCellDynamics(count) and
    Stream(count)); then every pair's average into its coarse block, gated by
    the coarse level (AverageCommunicate(count)); then every pair's
    interpolation into its fine block, gated by the fine level
    (InterpolateCommunicate(count)).

    patches: (parent, origin, m) per patch, in creation order: the index of
    the parent block (0 is the cavity), the parent cell where the patch
    starts (per axis, the parent's local index, as RefinedCavity's origin)
    and the patch extent in parent cells per axis. A patch block has 2m + 4
    cells per axis (two ghost layers) and the parent's cells inside the patch
    beyond its outer layer are COVERED. Each level's omega is
    RefineConverter::getOmegaF of its parent's (lbm/unit_converter.h:627).
    The cavity and its tasks are RefinedCavity's; a hierarchy of one patch
    is RefinedCavity."""

    def __init__(self, lattice, Ni, Nj, Nk, patches, RT=0.6, visc=0.01, u_phys=0.05, dtype=np.float32):
        q = 19 if lattice == "d3q19" else 9
        d = 3 if q == 19 else 2
        self.q, self.d, self.dtype = q, d, dtype
        omega, ulat = convert_from_rt(np.float32(RT), np.float32(visc), np.float32(u_phys))
        if dtype == np.float64:
            omega = np.float64(1.0 / RT)
        self.ulat = ulat
        cflags = cavity_flags(Ni, Nj, Nk if d == 3 else 0)
        dims = [(Ni + 2, Nj + 2, (Nk + 2) if d == 3 else 1)]
        flags = [cflags.copy()]
        levels, omegas, self.patches = [0], [omega], []
        for parent, origin, m in patches:
            o = [int(v) for v in (list(origin) + [0, 0])[:3]]
            mm = [int(v) for v in (list(m) + [1, 1])[:3]] if not np.isscalar(m) else [int(m)] * d + [1] * (3 - d)
            pd = dims[parent]
            f3 = flags[parent].reshape(pd[::-1])
            inner = tuple(slice(o[a] + 1, o[a] + mm[a] - 1) for a in range(d))
            if d == 3:
                f3[inner[::-1]] = COVERED
            else:
                f3[0][inner[::-1]] = COVERED
            fd = tuple(2 * mm[a] + 4 if a < d else 1 for a in range(3))
            dims.append(fd)
            flags.append(np.full(int(np.prod(fd)), AABB, dtype=np.uint8))
            levels.append(levels[parent] + 1)
            om_p = omegas[parent]
            if dtype == np.float64:
                omegas.append(np.float64(1.0) / (np.float64(2.0) / om_p - np.float64(0.5)))
            else:
                omegas.append(np.float32(1) / (np.float32(2) / np.float32(om_p) - np.float32(0.5)))
            self.patches.append((parent, len(dims) - 1, o))
        self.levels, self.omegas, self.dims = levels, omegas, dims
        self.max_level = max(levels)
        tasks = [(AABB, DYN_BGK_WRITE), (BB | BBMW, DYN_BGK_FIELD)]
        bcs = [(BB, VOID, HWBC_NORMAL), (BBMW, VOID, HWBC_MOVING)]
        w = weights(q, dtype)
        self.blocks, self.pops = [], []
        for k, (dm, fl) in enumerate(zip(dims, flags)):
            N = fl.size
            u = np.zeros((d, N))
            if k == 0:
                u[0, fl == BBMW] = float(ulat)
            self.blocks.append(Lattice(q, dm, fl, tasks, bcs if k == 0 else [], omega=float(omegas[k]),
                                       rho_field=np.ones(N), u_field=u))
            self.pops.append(np.ascontiguousarray(np.repeat(w[:, None], N, axis=1).astype(dtype)))
        self.count = 0

    def due(self, level, count):
        return count % (1 << (self.max_level - level)) == 0

    def run(self, n):
        for _ in range(n):
            c = self.count
            for k, b in enumerate(self.blocks):
                if self.due(self.levels[k], c):
                    b.run(self.pops[k], 1)
            for pa, ch, o in self.patches:
                if self.due(self.levels[pa], c):
                    refine_average(self.blocks[pa], self.pops[pa], self.blocks[ch], self.pops[ch], o)
            for pa, ch, o in self.patches:
                if self.due(self.levels[ch], c):
                    refine_interpolate(self.blocks[pa], self.pops[pa], self.blocks[ch], self.pops[ch], o)
            self.count += 1


def cavity_flags(Ni, Nj, Nk=0):
    n = (Ni + 2) * (Nj + 2) * ((Nk + 2) if Nk > 0 else 1)
    out = np.empty(n, dtype=np.uint8)
    lib().flbo_cavity_flags(Ni, Nj, Nk, _ptr(out, C.c_uint8))
    return out


def convert_from_rt(RT, visc, u_phys):
    om, ul = C.c_float(), C.c_float()
    lib().flbo_convert_from_rt(RT, visc, u_phys, C.byref(om), C.byref(ul))
    return np.float32(om.value), np.float32(ul.value)


def weights(q, dtype=np.float32):
    one = dtype(1)
    if q == 19:
        return np.array([one / dtype(3)] + [one / dtype(18)] * 6 + [one / dtype(36)] * 12, dtype=dtype)
    return np.array([dtype(4) / dtype(9)] + [one / dtype(9)] * 4 + [one / dtype(36)] * 4, dtype=dtype)


def cavity_case(mode, Ni, Nj, Nk, RT=0.53, visc=0.01, u_phys=0.1, dtype=np.float32, cs=0.17, rates=None):
    """The reference cavity driver as an oracle Lattice plus its initial pops.

    mode 'fw'      : benchmarks/cavity3d/cavity3d.cpp:170-182 (wall cells swap).
    mode 'hw'      : every non-void cell BGK<rhoU>; halfway normal / moving wall.
    mode 'hwfield' : examples/couette2d/couette2d.cpp:183-216 (walls BGK<useFieldrhoU>).
    mode 'mrt'     : fullway walls, fluid MRT_Feq_RhoU (D2Q9 in FreeLB).
    mode 'smag'    : fullway walls, fluid SmagorinskyBGK with C_s = cs.
    mode 'force'   : like 'hw' with BGKForce<forcerhoU, SecondOrder, ConstForce> and the
                     constant force density FORCE (float values).
    mode 'mrt_hw'  : every non-void cell MRT; halfway normal / moving wall.
    mode 'smag_mrt', 'smag_mrt_hw' : as 'mrt', 'mrt_hw' with MRT + Smagorinsky (C_s = cs).
    rates          : MRT rates (NaN: the shear rate), None for the defaults.
    """
    q = 19 if Nk > 0 else 9
    flags = cavity_flags(Ni, Nj, Nk)
    dims = (Ni + 2, Nj + 2, (Nk + 2) if Nk > 0 else 1)
    omega, ulat = convert_from_rt(np.float32(RT), np.float32(visc), np.float32(u_phys))
    d = 3 if q == 19 else 2
    N = flags.size
    rho_field = np.ones(N)
    u_field = np.zeros((d, N))
    u_field[0, flags == BBMW] = float(ulat)
    if mode == "fw":
        tasks = [(AABB, DYN_BGK), (BB, DYN_BOUNCEBACK), (BBMW, DYN_BOUNCEBACK_MW)]
        bcs = []
    elif mode == "hw":
        tasks = [(AABB | BB | BBMW, DYN_BGK)]
        bcs = [(BB, VOID, HWBC_NORMAL), (BBMW, VOID, HWBC_MOVING)]
    elif mode == "hwfield":
        tasks = [(AABB, DYN_BGK), (BB | BBMW, DYN_BGK_FIELD)]
        bcs = [(BB, VOID, HWBC_NORMAL), (BBMW, VOID, HWBC_MOVING)]
    elif mode == "mrt":
        tasks = [(AABB, DYN_MRT), (BB, DYN_BOUNCEBACK), (BBMW, DYN_BOUNCEBACK_MW)]
        bcs = []
    elif mode == "smag":
        tasks = [(AABB, DYN_SMAG_BGK, float(np.float32(cs))), (BB, DYN_BOUNCEBACK),
                 (BBMW, DYN_BOUNCEBACK_MW)]
        bcs = []
    elif mode == "force":
        tasks = [(AABB | BB | BBMW, DYN_BGK_FORCE)]
        bcs = [(BB, VOID, HWBC_NORMAL), (BBMW, VOID, HWBC_MOVING)]
    elif mode == "mrt_hw":
        tasks = [(AABB | BB | BBMW, DYN_MRT)]
        bcs = [(BB, VOID, HWBC_NORMAL), (BBMW, VOID, HWBC_MOVING)]
    elif mode == "smag_mrt":
        tasks = [(AABB, DYN_SMAG_MRT, float(np.float32(cs))), (BB, DYN_BOUNCEBACK),
                 (BBMW, DYN_BOUNCEBACK_MW)]
        bcs = []
    elif mode == "smag_mrt_hw":
        tasks = [(AABB | BB | BBMW, DYN_SMAG_MRT, float(np.float32(cs)))]
        bcs = [(BB, VOID, HWBC_NORMAL), (BBMW, VOID, HWBC_MOVING)]
    else:
        raise ValueError(mode)
    if dtype == np.float64:
        om = 1.0 / RT
        ul = float(u_phys) / (1.0 / ((RT - 0.5) * (1.0 / 3.0) / visc))
        u_field[0, flags == BBMW] = ul
    else:
        om = float(omega)
    force = [float(v) for v in FORCE[:d]] if mode == "force" else (0.0, 0.0, 0.0)
    lat = Lattice(q, dims, flags, tasks, bcs, omega=om, rho_field=rho_field, u_field=u_field,
                  force=force, mrt_rates=rates)
    w = weights(q, dtype)
    pops = np.repeat(w[:, None], N, axis=1).astype(dtype)
    pops = np.ascontiguousarray(pops)
    return lat, pops
