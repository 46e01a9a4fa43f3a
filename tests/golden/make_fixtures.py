#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ by running FreeLB itself.

FreeLB is built from its own sources (header-only C++, /root/reference/src) by
oracle/Makefile into oracle/_ref/cavity_ref, a driver that sets up the
lid-driven cavity exactly as benchmarks/cavity3d/cavity3d.cpp does and dumps
the populations, the flag field and FreeLB's own moment::rhoU read-out.
This script only runs that binary and stores its outputs; it needs the
reference sources, so it runs in the build container, not on the GPU box.

    python tests/golden/make_fixtures.py [--only MODE]   (MODE: write only that mode's small fixtures,
                                                          "refined" for the refined-cavity ones,
                                                          "vtk" for FreeLB's VTK files,
                                                          "natconv" for the coupled-lattice ones,
                                                          "shipped" for the as-shipped build's one)
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "cavity_ref")

# (lattice, mode, Ni, Nj, Nk, steps): full-population fixtures
SMALL = [
    ("d3q19", "fw", 8, 7, 6, 40), ("d3q19", "hw", 8, 7, 6, 40), ("d3q19", "hwfield", 8, 7, 6, 40),
    ("d3q19", "hw", 8, 7, 6, 39),
    ("d2q9", "fw", 16, 13, 1, 100), ("d2q9", "hw", 16, 13, 1, 100), ("d2q9", "hwfield", 16, 13, 1, 100),
    ("d2q9", "mrt", 16, 13, 1, 100),
    ("d2q9", "smag", 16, 13, 1, 100), ("d3q19", "smag", 8, 7, 6, 40),
    ("d3q19", "force", 8, 7, 6, 40), ("d2q9", "force", 16, 13, 1, 100),
]
CS = "0.17"  # Smagorinsky constant of the smag fixtures
# the 100^3 cavity, 1000 steps: hashes and a mid-plane slice
LARGE = [("d3q19", "fw", 100, 100, 100, 1000), ("d3q19", "hw", 100, 100, 100, 1000)]
RT, VISC, U = "0.53", "0.01", "0.1"
# two-level refined cavity (oracle/_ref/refine_ref, examples/cavref2d/cavref2d.cpp's loop):
# (lattice, Ni=Nj(=Nk), blockcell, p0, p1, counts); FreeLB's 3D refinement builder
# writes past its fine block, so only D2Q9 is pinned this way
REFINED = [("d2q9", 16, 4, 1, 3, 41), ("d2q9", 24, 4, 2, 4, 34)]
# three levels: block cells in [p0, p1) refined once, those in [q0, q1) twice
# (lattice, Ni, blockcell, p0, p1, q0, q1, counts)
REFINED3 = [("d2q9", 24, 4, 1, 5, 2, 4, 37), ("d2q9", 32, 4, 1, 7, 3, 5, 44)]
REF_RT, REF_VISC, REF_U = "0.6", "0.01", "0.05"
EXE_REF = os.path.join(ROOT, "oracle", "_ref", "refine_ref")
# FreeLB's own VTK output (oracle/_ref/vtk_ref, vtmo::vtmWriter) of the halfway cavity after the
# rhoU<CELL, true> read-out: (lattice, Ni, Nj, Nk, steps, overlap threshold, name)
VTK = [("d3q19", 8, 7, 6, 40, 0, "cav3"), ("d2q9", 16, 13, 1, 100, -1, "cav2")]
EXE_VTK = os.path.join(ROOT, "oracle", "_ref", "vtk_ref")
# coupled flow + temperature lattices, a differentially heated cavity (oracle/_ref/natconv_ref,
# the coupling loop of examples/cazsblock2d/cazsblock2d.cpp): (lattice, Ni, Nj, Nk, steps, RT_flow,
# RT_temp, gbeta, T_hot, T_cold, T_init)
NATCONV = [("d2q9", 24, 16, 1, 300, "0.6", "0.7", "0.001", "1.5", "0.5", "1.0"),
           ("d3q19", 10, 8, 6, 120, "0.6", "0.8", "0.002", "1.5", "0.5", "1.0")]
EXE_NC = os.path.join(ROOT, "oracle", "_ref", "natconv_ref")
# The same harness built with FreeLB's benchmark flags (floating-point
# contraction on, -march=native): the halfway 100^3 cavity after 1000 steps,
# rho / u on a sample of cells (two planes, a strided lattice of cells and the
# cells where the two builds differ most), plus the full-field differences
# between that build and the contraction-off one the bit-exact fixtures use.
EXE_SHIPPED = os.path.join(ROOT, "oracle", "_ref", "cavity_ref_shipped")
SHIPPED = ("d3q19", "hw", 100, 100, 100, 1000)


def read_dump(path):
    b = open(path, "rb").read()
    assert b[:8] == b"FLBREF01"
    d, q, nx, ny, nz, steps = (int(v) for v in np.frombuffer(b[8:32], dtype=np.int32))
    omega, uw = np.frombuffer(b[32:40], dtype=np.float32)
    n = nx * ny * nz
    off = 40
    flags = np.frombuffer(b[off:off + n], dtype=np.uint8).copy()
    off += n
    pops = np.frombuffer(b[off:off + 4 * q * n], dtype=np.float32).reshape(q, n).copy()
    off += 4 * q * n
    rhou = np.frombuffer(b[off:off + 4 * (d + 1) * n], dtype=np.float32).reshape(d + 1, n).copy()
    return dict(d=d, q=q, dims=(nx, ny, nz), steps=steps, omega=np.float32(omega),
                uw=np.float32(uw), flags=flags, pops=pops, rhou=rhou)


def read_refined(path):
    """FLBREF02: every block of the refined run (level, overlap, mesh, omega,
    voxel size, min corner, flags, pops, RHO/VELOCITY fields)."""
    b = open(path, "rb").read()
    assert b[:8] == b"FLBREF02"
    d, q, nb, steps = (int(v) for v in np.frombuffer(b[8:24], dtype=np.int32))
    uw = np.frombuffer(b[24:28], dtype=np.float32)[0]
    off = 28
    blocks = []
    for _ in range(nb):
        lev, ov, nx, ny, nz = (int(v) for v in np.frombuffer(b[off:off + 20], dtype=np.int32))
        off += 20
        om, vox, mx, my, mz = np.frombuffer(b[off:off + 20], dtype=np.float32)
        off += 20
        n = nx * ny * nz
        flags = np.frombuffer(b[off:off + n], dtype=np.uint8).copy()
        off += n
        pops = np.frombuffer(b[off:off + 4 * q * n], dtype=np.float32).reshape(q, n).copy()
        off += 4 * q * n
        fields = np.frombuffer(b[off:off + 4 * (d + 1) * n], dtype=np.float32).reshape(d + 1, n).copy()
        off += 4 * (d + 1) * n
        blocks.append(dict(level=lev, overlap=ov, dims=(nx, ny, nz), omega=om, vox=vox,
                           min=np.array([mx, my, mz], dtype=np.float32), flags=flags, pops=pops, fields=fields))
    return dict(d=d, q=q, steps=steps, uw=uw, blocks=blocks)


def read_natconv(path):
    """FLBNC001: flags, the initial and final flow / temperature populations,
    the final RHO, VELOCITY, TEMP and SCALARFORCE fields."""
    b = open(path, "rb").read()
    assert b[:8] == b"FLBNC001"
    d, qf, qt, nx, ny, nz, steps = (int(v) for v in np.frombuffer(b[8:36], dtype=np.int32))
    prm = np.frombuffer(b[36:60], dtype=np.float32).copy()
    n = nx * ny * nz
    off = 60
    flags = np.frombuffer(b[off:off + n], dtype=np.uint8).copy()
    off += n

    def take(k):
        nonlocal off
        a = np.frombuffer(b[off:off + 4 * k], dtype=np.float32).copy()
        off += 4 * k
        return a

    out = dict(d=d, qf=qf, qt=qt, dims=np.array([nx, ny, nz]), steps=steps, params=prm, flags=flags)
    out["flow0"] = take(qf * n).reshape(qf, n)
    out["temp0"] = take(qt * n).reshape(qt, n)
    out["T0"] = take(n)
    out["flow"] = take(qf * n).reshape(qf, n)
    out["temp"] = take(qt * n).reshape(qt, n)
    out["rho"] = take(n)
    out["u"] = take(d * n).reshape(d, n)
    out["T"] = take(n)
    out["sforce"] = take(n)
    assert off == len(b)
    return out


def run_refined(lat, n, cell, p0, p1, counts, tmp, q0=None, q1=None):
    out = os.path.join(tmp, f"refine_{lat}_{n}_{cell}_{p0}_{p1}_{counts}.bin")
    nk = n if lat == "d3q19" else 1
    extra = [] if q0 is None else [str(q0), str(q1)]
    subprocess.run([EXE_REF, lat, str(n), str(n), str(nk), str(cell), str(p0), str(p1), str(counts), REF_RT,
                    REF_VISC, REF_U, out, *extra], check=True, capture_output=True)
    return read_refined(out)


def run(lat, mode, ni, nj, nk, steps, tmp):
    out = os.path.join(tmp, f"{lat}_{mode}_{ni}_{nj}_{nk}_{steps}.bin")
    env = dict(os.environ, OMP_NUM_THREADS=os.environ.get("OMP_NUM_THREADS", "8"))
    args = [EXE, lat, mode, str(ni), str(nj), str(nk), str(steps), RT, VISC, U, out]
    if mode == "smag":
        args.append(CS)
    subprocess.run(args, check=True, capture_output=True, env=env)
    return read_dump(out)


def shipped_errors(ref, got, nonvoid):
    """rho: max |d rho| / |rho|; u: max |d u| / max |u| over the cells, and the
    pointwise form max |d u| / max(|u|, 1e-3 max |u|)."""
    r, g = ref[:, nonvoid].astype(np.float64), got[:, nonvoid].astype(np.float64)
    umax = float(np.abs(r[1:]).max())
    du = np.abs(g[1:] - r[1:])
    return {"rho_max_rel": float((np.abs(g[0] - r[0]) / np.abs(r[0])).max()),
            "u_max_abs": float(du.max()), "u_max_over_umax": float(du.max() / umax),
            "u_pointwise_floor1e-3": float((du / np.maximum(np.abs(r[1:]), 1e-3 * umax)).max()),
            "u_max": umax}


def write_shipped(tmp):
    if not os.path.exists(EXE_SHIPPED):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "_ref/cavity_ref_shipped"], check=True)
    lat, mode, ni, nj, nk, steps = SHIPPED
    off = run(lat, mode, ni, nj, nk, steps, tmp)
    args = [EXE_SHIPPED, lat, mode, str(ni), str(nj), str(nk), str(steps), RT, VISC, U,
            os.path.join(tmp, "shipped.bin")]
    env = dict(os.environ, OMP_NUM_THREADS=os.environ.get("OMP_NUM_THREADS", "8"))
    subprocess.run(args, check=True, capture_output=True, env=env)
    shp = read_dump(os.path.join(tmp, "shipped.bin"))
    nx, ny, nz = shp["dims"]
    nonvoid = shp["flags"] != 1
    # sample: the mid-plane y = ny // 2, the lid's fluid layer, every 5th cell
    # per axis, and the 2000 cells whose u differs most between the builds
    ids = np.arange(nx * ny * nz).reshape(nz, ny, nx)
    pick = [ids[:, ny // 2, :].ravel(), ids[nz - 2, :, :].ravel(), ids[1:-1:5, 1:-1:5, 1:-1:5].ravel()]
    du = np.abs(shp["rhou"][1:].astype(np.float64) - off["rhou"][1:]).max(axis=0)
    du[~nonvoid] = -1.0
    pick.append(np.argsort(du)[-2000:])
    cells = np.unique(np.concatenate(pick))
    cells = cells[nonvoid[cells]].astype(np.int32)
    stats = shipped_errors(off["rhou"], shp["rhou"], nonvoid)
    name = f"cavity_{lat}_{mode}_{ni}x{nj}x{nk}_{steps}_shipped.npz"
    np.savez_compressed(os.path.join(HERE, name), cells=cells, rhou=shp["rhou"][:, cells],
                        dims=np.array([nx, ny, nz]), steps=steps,
                        stats_vs_contraction_off=json.dumps(stats))
    print("wrote", name, stats, flush=True)


def main():
    only = sys.argv[sys.argv.index("--only") + 1] if "--only" in sys.argv else None
    if not os.path.exists(EXE):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    with tempfile.TemporaryDirectory() as tmp:
        if only in (None, "vtk"):
            if not os.path.exists(EXE_VTK):
                subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "_ref/vtk_ref"], check=True)
            outdir = os.path.join(HERE, "vtk")
            os.makedirs(outdir, exist_ok=True)
            for lat, ni, nj, nk, steps, olth, name in VTK:
                work = os.path.join(tmp, name)
                os.makedirs(work)
                subprocess.run([EXE_VTK, lat, str(ni), str(nj), str(nk), str(steps), str(olth), name], cwd=work,
                               check=True, capture_output=True)
                src = os.path.join(work, "vtkoutput", "vtidata")
                for fn in sorted(os.listdir(src)):
                    with open(os.path.join(src, fn), "rb") as a, open(os.path.join(outdir, fn), "wb") as b:
                        b.write(a.read())
                    print("wrote vtk/" + fn, flush=True)
            if only:
                return
        if only in (None, "shipped"):
            write_shipped(tmp)
            if only:
                return
        if only in (None, "natconv"):
            if not os.path.exists(EXE_NC):
                subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "_ref/natconv_ref"], check=True)
            for case in NATCONV:
                lat, ni, nj, nk, steps = case[:5]
                out = os.path.join(tmp, f"natconv_{lat}.bin")
                subprocess.run([EXE_NC, lat, str(ni), str(nj), str(nk), str(steps), *case[5:], out], check=True,
                               capture_output=True)
                r = read_natconv(out)
                name = f"natconv_{lat}_{ni}x{nj}x{nk}_{steps}.npz"
                np.savez_compressed(os.path.join(HERE, name), **{k: np.asarray(v) for k, v in r.items()})
                print("wrote", name, flush=True)
            if only:
                return
        if only in (None, "refined"):
            if not os.path.exists(EXE_REF):
                subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "_ref/refine_ref"], check=True)
            for lat, n, cell, p0, p1, counts in REFINED:
                r = run_refined(lat, n, cell, p0, p1, counts, tmp)
                arrays = {"case": np.array([n, cell, p0, p1, counts]), "uw": r["uw"]}
                for k, blk in enumerate(r["blocks"]):
                    for key in ("level", "overlap", "dims", "omega", "min", "flags", "pops", "fields"):
                        arrays[f"b{k}_{key}"] = np.asarray(blk[key])
                arrays["nblocks"] = np.array(len(r["blocks"]))
                name = f"refined_{lat}_{n}_cell{cell}_p{p0}{p1}_{counts}.npz"
                np.savez_compressed(os.path.join(HERE, name), **arrays)
                print("wrote", name, flush=True)
            for lat, n, cell, p0, p1, q0, q1, counts in REFINED3:
                r = run_refined(lat, n, cell, p0, p1, counts, tmp, q0, q1)
                arrays = {"case": np.array([n, cell, p0, p1, counts, q0, q1]), "uw": r["uw"]}
                for k, blk in enumerate(r["blocks"]):
                    for key in ("level", "overlap", "dims", "omega", "min", "flags", "pops", "fields"):
                        arrays[f"b{k}_{key}"] = np.asarray(blk[key])
                arrays["nblocks"] = np.array(len(r["blocks"]))
                name = f"refined3_{lat}_{n}_cell{cell}_p{p0}{p1}_q{q0}{q1}_{counts}.npz"
                np.savez_compressed(os.path.join(HERE, name), **arrays)
                print("wrote", name, flush=True)
            if only:
                return
        for lat, mode, ni, nj, nk, steps in SMALL:
            if only and mode != only:
                continue
            r = run(lat, mode, ni, nj, nk, steps, tmp)
            name = f"cavity_{lat}_{mode}_{ni}x{nj}x{nk}_{steps}.npz"
            np.savez_compressed(os.path.join(HERE, name), flags=r["flags"], pops=r["pops"],
                                rhou=r["rhou"], dims=np.array(r["dims"]), steps=r["steps"],
                                omega=r["omega"], uw=r["uw"])
            print("wrote", name, flush=True)
        if only:
            return
        meta = {}
        for lat, mode, ni, nj, nk, steps in LARGE:
            r = run(lat, mode, ni, nj, nk, steps, tmp)
            nx, ny, nz = r["dims"]
            nonvoid = r["flags"] != 1
            key = f"{lat}_{mode}_{ni}x{nj}x{nk}_{steps}"
            meta[key] = {
                "dims": [nx, ny, nz],
                "sha256_pops_all_cells": hashlib.sha256(r["pops"].tobytes()).hexdigest(),
                "sha256_rhou_nonvoid": hashlib.sha256(r["rhou"][:, nonvoid].tobytes()).hexdigest(),
                "omega_bits": int(r["omega"].view(np.uint32)),
                "uw_bits": int(r["uw"].view(np.uint32)),
            }
            mid = ny // 2
            sl = r["rhou"].reshape(r["d"] + 1, nz, ny, nx)[:, :, mid, :]
            np.savez_compressed(os.path.join(HERE, f"cavity_{key}_midplane.npz"), rhou_y=sl,
                                y=mid)
            print("wrote", key, flush=True)
        with open(os.path.join(HERE, "cavity_large.json"), "w") as fh:
            json.dump(meta, fh, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
