#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ by running FreeLB itself.

FreeLB is built from its own sources (header-only C++, /root/reference/src) by
oracle/Makefile into oracle/_ref/cavity_ref, a driver that sets up the
lid-driven cavity exactly as benchmarks/cavity3d/cavity3d.cpp does and dumps
the populations, the flag field and FreeLB's own moment::rhoU read-out.
This script only runs that binary and stores its outputs; it needs the
reference sources, so it runs in the build container, not on the GPU box.

    python tests/golden/make_fixtures.py [--only MODE]   (MODE: write only that mode's small fixtures)
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


def run(lat, mode, ni, nj, nk, steps, tmp):
    out = os.path.join(tmp, f"{lat}_{mode}_{ni}_{nj}_{nk}_{steps}.bin")
    env = dict(os.environ, OMP_NUM_THREADS=os.environ.get("OMP_NUM_THREADS", "8"))
    args = [EXE, lat, mode, str(ni), str(nj), str(nk), str(steps), RT, VISC, U, out]
    if mode == "smag":
        args.append(CS)
    subprocess.run(args, check=True, capture_output=True, env=env)
    return read_dump(out)


def main():
    only = sys.argv[sys.argv.index("--only") + 1] if "--only" in sys.argv else None
    if not os.path.exists(EXE):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    with tempfile.TemporaryDirectory() as tmp:
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
