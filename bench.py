#!/usr/bin/env python3
"""Throughput of the fused D3Q19 lattice update on one MI355X.

Workload: the 256^3 lid-driven cavity (D3Q19, BGK, tau = 0.53, halfway
bounce-back on five walls, moving lid u = 0.1), populations resident in HBM.
One step = one fused collide-and-stream launch over the 256^3 fluid cells.
`value` is MLUPS of FP32 storage; the same workload with FP16 / BF16 shifted
storage is reported beside it, with each run's HBM roofline fraction
(152 B per cell update in FP32, 76 B in the shifted half formats).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure
N_EDGE = 256


class HipTimer:
    """hipEvent pair on a given stream (plain HIP runtime through ctypes)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.hip.hipDeviceSynchronize.argtypes = []
        self.a, self.b = C.c_void_p(), C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(self.a)) == 0
        assert self.hip.hipEventCreate(C.byref(self.b)) == 0

    def sync(self):
        assert self.hip.hipDeviceSynchronize() == 0

    def start(self, stream=None):
        assert self.hip.hipEventRecord(self.a, stream) == 0

    def stop(self, stream=None) -> float:
        assert self.hip.hipEventRecord(self.b, stream) == 0
        assert self.hip.hipEventSynchronize(self.b) == 0
        ms = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(ms), self.a, self.b) == 0
        return float(ms.value)


def make_block(storage, n=N_EDGE, collision="bgk"):
    from freelb_amd.cases import cavity_block
    return cavity_block(n, storage=storage, collision=collision,
                        tau=0.53 if collision == "bgk" else 0.5005)


def time_storage(storage, steps, warmup, timer):
    lat = make_block(storage)
    n_active = N_EDGE ** 3
    lat.step(warmup)
    lat.synchronize()
    timer.sync()
    t0 = time.perf_counter()
    timer.start()
    lat.step(steps)
    ms = timer.stop()
    lat.synchronize()
    wall = time.perf_counter() - t0
    mass = lat.mass_momentum()
    finite = bool(np.isfinite(mass).all())
    lat.close()
    bytes_per = 152 if storage == "f32" else (76 if storage in ("f16s", "bf16s", "f16c") else 304)
    ms_step = ms / steps
    mlups = n_active * steps / (ms / 1e3) / 1e6
    achieved = n_active * bytes_per / (ms_step / 1e3) / 1e9  # GB/s, algorithmic bytes per launch
    return {"storage": storage, "ms_per_step": ms_step, "mlups": mlups, "wall_s": wall,
            "achieved_gbs": achieved, "frac": achieved / HBM_PEAK_GBS, "bytes_per_cell": bytes_per,
            "finite": finite}


def mrt_les_speed(timer, storages=("f32", "f16s", "f16c"), n=512, steps=100, warmup=10):
    """Step rate of the 512^3 MRT+LES cavity (Re 5000, lid 0.05, D3Q19,
    d'Humieres MRT with the Smagorinsky shear rates): K timed steps after W
    untimed ones per storage kind, HIP events around the launches."""
    from freelb_amd.cases import cavity_block
    out = {}
    for s in storages:
        lat = cavity_block(n, storage=s, collision="mrt_les", tau=0.5 + 3 * 0.05 * n / 5000, u_lid=0.05)
        lat.step(warmup)
        lat.synchronize()
        timer.sync()
        timer.start()
        lat.step(steps)
        ms = timer.stop()
        finite = bool(np.isfinite(lat.mass_momentum()).all())
        lat.close()
        bpc = 152 if s == "f32" else (304 if s == "f64" else 76)
        gbs = n ** 3 * bpc * steps / (ms / 1e3) / 1e9
        out[s] = {"mlups": round(n ** 3 * steps / (ms / 1e3) / 1e6, 1), "ms_per_step": round(ms / steps, 5),
                  "achieved_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4), "finite": finite}
    return out


def _vorticity(rhou, n):
    """Central-difference curl of u over the cavity cells whose six
    neighbours are cavity cells; rhou is [4][(n+2)^3]."""
    u = [np.asarray(rhou[1 + d], dtype=np.float64).reshape(n + 2, n + 2, n + 2) for d in range(3)]
    ux, uy, uz = u  # indexed [z, y, x]
    c = slice(2, n)

    def d(f, axis):
        sl_p = [c, c, c]
        sl_m = [c, c, c]
        sl_p[axis] = slice(3, n + 1)
        sl_m[axis] = slice(1, n - 1)
        return 0.5 * (f[tuple(sl_p)] - f[tuple(sl_m)])
    # axis 0 = z, 1 = y, 2 = x
    return np.stack([d(uz, 1) - d(uy, 0), d(ux, 0) - d(uz, 2), d(uy, 2) - d(ux, 1)])


def precision_study(n=N_EDGE, steps=2000, collisions=("bgk", "mrt_les"),
                    storages=("f32", "f16s", "bf16s", "f16c"), seed=2024, u0_amp=1e-3, timer=None,
                    u_lid=0.1, tau=None):
    """Errors of FP32 / FP16-shift / BF16-shift / F16C storage against an FP64
    run of the same cavity (D3Q19, halfway walls, lid u_lid): rho and u from
    flb_moments, vorticity by central differences. Start: rho = 1, u i.i.d.
    U(-u0_amp, u0_amp) per component (numpy default_rng(seed)), f = feq.
    tau: BGK 0.53 and MRT+LES tau0 0.5005 unless given."""
    from freelb_amd.cases import VOID, cavity_block, cavity_flags
    live = cavity_flags(n) != VOID
    out = {"cells": n ** 3, "steps": steps, "seed": seed, "u0_amp": u0_amp, "u_lid": u_lid, "runs": {}}
    for coll in collisions:
        t = tau if tau is not None else (0.53 if coll == "bgk" else 0.5005)
        ref = None
        for s in ("f64",) + tuple(storages):
            lat = cavity_block(n, storage=s, collision=coll, tau=t, u0_amp=u0_amp, seed=seed, u_lid=u_lid)
            try:
                t0 = time.perf_counter()
                lat.step(steps)
                lat.synchronize()
                dt = time.perf_counter() - t0
                m = lat.moments()
            finally:
                lat.close()
            bpc = {"f64": 304, "f32": 152}.get(s, 76)
            rec = {"mlups": round(n ** 3 * steps / dt / 1e6, 1),
                   "hbm_frac": round(n ** 3 * steps * bpc / dt / (HBM_PEAK_GBS * 1e9), 4),
                   "finite": bool(np.isfinite(m[:, live]).all())}
            if s == "f64":
                ref = m
                vref = _vorticity(ref, n)
                rec["u_max"] = float(np.abs(ref[1:, live]).max())
            else:
                du = (m[1:, live].astype(np.float64) - ref[1:, live])
                rec["u_l2_rel"] = float(np.sqrt((du ** 2).sum() / (ref[1:, live] ** 2).sum()))
                rec["rho_max_abs"] = float(np.abs(m[0, live].astype(np.float64) - ref[0, live]).max())
                dv = _vorticity(m, n) - vref
                rec["vort_l2_rel"] = float(np.sqrt((dv ** 2).sum() / (vref ** 2).sum()))
            out["runs"][f"{coll}/{s}"] = rec
            print(f"# precision {coll}/{s}: {rec}", file=sys.stderr, flush=True)
    return out


def batch_specs(nblocks=512, seed=2025):
    """The batched-cavity workload: per block, edges drawn per axis from
    U{16..96}, tau from U[0.51, 0.9] and the lid speed from U[0.02, 0.1]
    (numpy default_rng(seed))."""
    rng = np.random.default_rng(seed)
    edges = rng.integers(16, 97, size=(nblocks, 3))
    taus = rng.uniform(0.51, 0.9, size=nblocks)
    lids = rng.uniform(0.02, 0.1, size=nblocks)
    return [dict(shape=tuple(int(v) for v in edges[k]), tau=float(taus[k]), u_lid=float(lids[k]))
            for k in range(nblocks)]


def batch_bench(timer, nblocks=512, seed=2025, graph_steps=500, storage="f32", loop_steps=20, eager=False):
    """512 independent D3Q19 BGK halfway cavities, edges drawn per axis from
    U{16..96}, tau from U[0.51, 0.9], lid speed from U[0.02, 0.1] (numpy
    default_rng(seed)), stepped together: one launch per step over the block
    table (autotuned shape), recorded as a hipGraph of graph_steps steps. The
    same blocks stepped with one launch per block and step (FreeLB's
    per-block manager loop) are timed beside it. eager: time graph_steps
    batch launches issued one by one instead of the graph (for kernel
    profilers that do not follow graph launches)."""
    from freelb_amd import BlockLatticeManager
    from freelb_amd.cases import cavity_block
    specs = batch_specs(nblocks, seed)
    t0 = time.perf_counter()
    blocks = [cavity_block(sp["shape"], storage=storage, tau=sp["tau"], u_lid=sp["u_lid"]) for sp in specs]
    setup_s = time.perf_counter() - t0
    cells = int(sum(np.prod(sp["shape"]) for sp in specs))
    try:
        # one launch per block and step
        for b in blocks:
            b.step(2)
        timer.sync()
        timer.start()
        for _ in range(loop_steps // 2):
            for b in blocks:
                b.step(2)
        loop_ms = timer.stop() / loop_steps
        with BlockLatticeManager(blocks) as m:
            tuned = m.autotune(10)
            if eager:
                m.step(2)
                m.synchronize()
                timer.start()
                m.step(graph_steps)
                graph_ms = timer.stop() / graph_steps
            else:
                m.capture(graph_steps)
                m.replay()
                m.synchronize()
                timer.start()
                m.replay()
                graph_ms = timer.stop() / graph_steps
            m.synchronize()
            mass = np.array([b.mass_momentum()[0] for b in blocks])
    finally:
        for b in blocks:
            b.close()
    # blocks pairing tau near 0.51 with a lid near 0.1 are unstable under BGK
    # (the oracle diverges on them too); the rate counts every block's cells
    diverged = [k for k in range(nblocks) if not np.isfinite(mass[k])]
    bytes_per = 152 if storage == "f32" else 76
    how = f"{graph_steps} eager batch launches" if eager else f"hipGraph of {graph_steps} steps"
    return {"workload": f"{nblocks} cavities, edges U{{16..96}}^3 (seed {seed}), tau U[0.51,0.9], "
                        f"lid U[0.02,0.1], {how}",
            "blocks": nblocks, "cells": cells, "storage": storage,
            "mlups": round(cells / (graph_ms / 1e3) / 1e6, 1),
            "ms_per_step": round(graph_ms, 5),
            "achieved_gbs": round(cells * bytes_per / (graph_ms / 1e3) / 1e9, 1),
            "per_block_launch_mlups": round(cells / (loop_ms / 1e3) / 1e6, 1),
            "launch": tuned, "setup_s": round(setup_s, 1),
            "finite_blocks": nblocks - len(diverged),
            "diverged_blocks": [(k, round(specs[k]["tau"], 4), round(specs[k]["u_lid"], 3)) for k in diverged]}


# algorithmic HBM bytes per cell of the FP32 adjoint kernels, for boxes
# without solid cells (their link words come from the cell's place; a box
# with solid cells also reads a 4-byte link word per cell and kernel):
# - forward push: reads and writes 19 populations, reads s (4 B);
# - moment record (rho - 1, u) a recompute step, or a forward step with
#   checkpoint_every 0, also writes: 16 B;
# - moment record alone (the last state of each recomputed segment): reads
#   19 populations and s, writes 16 B;
# - adjoint step: reads the moment record and 19 adjoint values, writes 19
#   adjoint values, reads s and updates dJ/ds (8 B).
ADJ_FWD_BYTES = 19 * 4 * 2 + 4
ADJ_MOM_BYTES = 4 * 4
ADJ_MOMONLY_BYTES = 19 * 4 + 4 + 4 * 4
ADJ_BWD_BYTES = 4 * 4 + 19 * 4 * 2 + 4 + 8


def _adjoint_rates(cells, T, K, fwd_ms, bwd_ms):
    """Cell updates per second of the forward pass, of the adjoint pass
    (T adjoint steps plus the T - ceil(T/K) recomputed forward steps and the
    ceil(T/K) moment records of segment ends) and of the whole gradient, with
    the bandwidth each pass reaches."""
    nseg = -(-T // K) if K > 0 else 0
    rec = T - nseg if K > 0 else 0
    fwd_b = cells * T * (ADJ_FWD_BYTES + (ADJ_MOM_BYTES if K == 0 else 0))
    bwd_b = cells * (T * ADJ_BWD_BYTES + rec * (ADJ_FWD_BYTES + ADJ_MOM_BYTES) + nseg * ADJ_MOMONLY_BYTES)
    return {"forward_mlups": round(cells * T / fwd_ms / 1e3, 1),
            "adjoint_mlups": round(cells * T / bwd_ms / 1e3, 1),
            "gradient_mlups": round(cells * T / (fwd_ms + bwd_ms) / 1e3, 1),
            "forward_ms": round(fwd_ms, 2), "backward_ms": round(bwd_ms, 2),
            "forward_gbs": round(fwd_b / fwd_ms / 1e6, 1), "backward_gbs": round(bwd_b / bwd_ms / 1e6, 1),
            "forward_frac": round(fwd_b / fwd_ms / 1e6 / HBM_PEAK_GBS, 4),
            "backward_frac": round(bwd_b / bwd_ms / 1e6 / HBM_PEAK_GBS, 4),
            "recomputed_steps": rec}


def adjoint_bench(T=800, K=28):
    """The differentiable channel, 192 x 96 x 96 cells, periodic in x, walls in
    y and z, body force 1e-6 along x, tau 0.6, Brinkman fraction from the
    seed-7 smoothed field: FP32 forward of T steps with a checkpoint every K
    steps, then the adjoint pass (second run timed; the first allocates)."""
    from freelb_amd.cases import channel_adjoint
    adj, s = channel_adjoint(storage="f32")
    with adj:
        adj.forward(T, K)
        adj.backward()
        J = adj.forward(T, K)
        g = adj.backward()
        fwd_ms, bwd_ms = adj.timing()
    rec = f"checkpoint every {K}" if K > 0 else "moment record of every state (checkpoint_every 0), no recompute"
    out = {"workload": f"channel 192x96x96, T {T}, {rec}, FP32", "cells": int(s.size),
           "J": J, "grad_finite": bool(np.isfinite(g).all())}
    out.update(_adjoint_rates(s.size, T, K, fwd_ms, bwd_ms))
    return out


def adjoint_batch_bench(n=64, T=400, seed=11, fp64_check=False, K=None):
    """n channels with seeded extents uniform in [64,32,32]..[128,64,64]
    (numpy default_rng(seed)), each with its own seeded Brinkman field,
    stepped as one batch (one launch per step over all boxes), T steps with
    a checkpoint every round(sqrt(T)) steps (or every K; K = 0 records every
    state's moments and recomputes nothing), then the batched adjoint. With
    fp64_check the same batch runs in FP64 and the FP32 objective and
    gradient errors against it are reported."""
    from freelb_amd.cases import channel_batch, channel_dims
    K = int(round(T ** 0.5)) if K is None else int(K)
    dims = channel_dims(n, seed=seed)
    cells = int(sum(int(np.prod(d)) for d in dims))
    t0 = time.perf_counter()
    b, mem, _ = channel_batch(dims, storage="f32", seed=seed)
    setup_s = time.perf_counter() - t0
    try:
        b.forward(T, K)
        b.backward()
        J32 = b.forward(T, K)
        g32 = b.backward()
        fwd_ms, bwd_ms = b.timing()
    finally:
        b.close()
        for m in mem:
            m.close()
    rec = f"checkpoint every {K}" if K > 0 else "moment record of every state (checkpoint_every 0), no recompute"
    out = {"workload": f"{n} channels, extents U[64..128]xU[32..64]xU[32..64] (seed {seed}), T {T}, "
                       f"{rec}, FP32, one launch per step", "boxes": n, "cells": cells,
           "setup_s": round(setup_s, 1), "grad_finite": bool(all(np.isfinite(g).all() for g in g32))}
    out.update(_adjoint_rates(cells, T, K, fwd_ms, bwd_ms))
    if fp64_check:
        b, mem, _ = channel_batch(dims, storage="f64", seed=seed)
        try:
            J64 = b.forward(T, K)
            g64 = b.backward()
        finally:
            b.close()
            for m in mem:
                m.close()
        out["J_max_rel_err_vs_fp64"] = float(np.max(np.abs(J32 - J64) / np.abs(J64)))
        out["grad_max_err_vs_fp64"] = float(max(np.abs(a - c).max() / np.abs(c).max() for a, c in zip(g32, g64)))
        out["grad_l2_rel_err_vs_fp64"] = float(max(np.linalg.norm(a - c) / np.linalg.norm(c)
                                                   for a, c in zip(g32, g64)))
    return out


def refined_bench(timer, n=128, patch=64, counts=200, storage="f32"):
    """The refined cavity: n^3 coarse D3Q19 cells with a centred patch^3 fine
    patch at twice the resolution, the loop of examples/cavref2d/cavref2d.cpp
    (fine step every count, coarse step on even counts, average and
    interpolation exchange) for `counts` counts after 2 untimed ones. Cell
    updates are the coarse cells stepped (every cell with a task: the cavity
    minus the covered part of the patch) plus the fine cells stepped (the
    patch and its inner ghost layer), each update reading and writing 19
    populations (152 B in FP32) and writing its moments to the fields (16 B)."""
    from freelb_amd.cases import refined_cavity
    coarse, fine, ref = refined_cavity(n, patch, storage)
    m = patch // 2
    try:
        ref.run(2)
        timer.sync()
        timer.start()
        ref.run(counts)
        ms = timer.stop()
        finite = bool(np.isfinite(coarse.mass_momentum()[0]) and np.isfinite(fine.mass_momentum()[0]))
        info = ref.info()
    finally:
        ref.close()
        coarse.close()
        fine.close()
    c_cells = n ** 3 - (m - 2) ** 3
    f_cells = (patch + 2) ** 3
    updates = c_cells * (counts // 2) + f_cells * counts
    bpu = (152 if storage == "f32" else 76) + 16
    return {"workload": f"refined cavity: {n}^3 coarse, {patch}^3 fine patch (2:1), {counts} fine steps, "
                        f"{storage}", "coarse_cells": c_cells, "fine_cells": f_cells,
            "exchange_items": info, "ms": round(ms, 3), "ms_per_coarse_step": round(2 * ms / counts, 4),
            "mlups": round(updates / (ms / 1e3) / 1e6, 1),
            "achieved_gbs": round(updates * bpu / (ms / 1e3) / 1e9, 1), "finite": finite}


def coupled_bench(timer, n=256, steps=100, warmup=4):
    """The differentially heated cavity (freelb_amd.cases.heated_cavity): a
    D3Q19 flow block with the buoyancy of a D3Q7 temperature block advected by
    its VELOCITY field, FreeLB's coupled-lattice loop (flow step, then
    temperature step) for `steps` iterations, FP32. One iteration updates
    every non-void cell of both lattices: per cell the flow step moves 152 B
    of populations, reads the temperature (4 B) and writes rho and u (16 B);
    the temperature step moves 56 B of populations, reads u (12 B) and writes
    the temperature (4 B): 244 B."""
    from freelb_amd.cases import heated_cavity
    flow, temp, cpl = heated_cavity(n, 3)
    try:
        cpl.step(warmup)
        timer.sync()
        timer.start()
        cpl.step(steps)
        ms = timer.stop()
        T = temp.download_fields()[0]
        u = flow.download_fields()[1:]
        finite = bool(np.isfinite(T).all() and np.isfinite(u).all())
        umax = float(np.abs(u).max())
    finally:
        temp.close()
        flow.close()
    cells = n ** 3
    bpu = 244
    return {"workload": f"heated cavity {n}^3: D3Q19 flow + D3Q7 temperature, buoyancy coupling, {steps} "
                        "iterations, f32", "cells": cells, "ms_per_iteration": round(ms / steps, 4),
            "mlups": round(cells * steps / (ms / 1e3) / 1e6, 1), "bytes_per_cell": bpu,
            "achieved_gbs": round(cells * steps * bpu / (ms / 1e3) / 1e9, 1), "u_max": round(umax, 6),
            "finite": finite}


def profiled_traffic(storage):
    """HBM bytes per launch of the step kernel for this workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (counters scaled by a
    copy kernel of known size in the same runs). None when absent."""
    path = os.path.join(ROOT, "profiles", "k_step_cell_d3q19_cavity256.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(rf"^traffic_bytes_per_launch {storage} = (\d+)", text, re.M)
    return int(m.group(1)) if m else None


def _fullway_reference(n, steps, threads):
    """FreeLB's own benchmarks/cavity3d driver (fullway walls), MLUPS over the
    fluid cells, or None when its binary is absent."""
    exe = os.path.join(ROOT, "oracle", "_ref", "cavity3d_bench")
    if not os.path.exists(exe):
        return None
    work = tempfile.mkdtemp(prefix="flb_cpu_")
    try:
        ini = (f"[workdir]\nworkdir_ = ./output\n[parallel]\nthread_num = {threads}\nblock_num = 1\n"
               f"[Mesh]\nNi = {n}\nNj = {n}\nNk = {n}\nCell_Len = 1\n"
               "[Physical_Property]\nrho_ref = 0.001\nKine_Visc = 0.01\n"
               "[Init_Conditions]\nU_Ini0 = 0\nU_Ini1 = 0\nU_Ini2 = 0\nU_Max = 0.1\n"
               "[Boundary_Conditions]\nVelo_Wall0 = 0.1\nVelo_Wall1 = 0\nVelo_Wall2 = 0\n"
               f"[LB]\nRT = 0.53\n[Simulation_Settings]\nTotalStep = {steps}\nOutputStep = 100000\n"
               "[tolerance]\ntol = 1e-4\n")
        with open(os.path.join(work, "cavity3d.ini"), "w") as fh:
            fh.write(ini)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close")
        out = subprocess.run([exe], cwd=work, env=env, capture_output=True, text=True, timeout=600).stdout
    finally:
        shutil.rmtree(work, ignore_errors=True)
    for line in out.splitlines():
        if line.strip().startswith("Average_MLUPs:"):  # utils/timer.h:93
            return float(line.split(":", 1)[1].split()[0]) * n ** 3 / (n + 2) ** 3
    return None


def cpu_baseline(n=N_EDGE, steps=200, warmup=5):
    """FreeLB itself on the host cores, on the workload of `value`: the 256^3
    halfway cavity (BGK<rhoU> on every cavity cell, bounceback::normal /
    movingwall managers after Stream, boundary/boundary.h:121-128), built
    from the reference sources with its benchmarks' flags
    (oracle/_ref/cavity_hw_bench), `steps` timed steps after `warmup`.
    FreeLB's published fullway benchmark (benchmarks/cavity3d) is timed
    beside it (fullway_value, 20 steps). Without the binaries, the C oracle
    (the same algorithm restated) is timed instead."""
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    exe = os.path.join(ROOT, "oracle", "_ref", "cavity_hw_bench")
    if os.path.exists(exe):
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close")
        out = subprocess.run([exe, "d3q19", "hw", str(n), str(n), str(n), str(steps), "0.53", "0.01", "0.1",
                              f"time:{warmup}"], env=env, capture_output=True, text=True, timeout=900).stdout
        # FreeLB prints its geometry summary first; the harness's result line
        # reads "MLUPS v cells N steps S seconds T threads n"
        lines = [ln.split() for ln in out.splitlines() if ln.startswith("MLUPS ")]
        fields = lines[-1] if lines else []
        if len(fields) >= 8:
            rec = {"value": float(fields[1]), "unit": "MLUPS", "cores": threads, "kind": "reference",
                   "sample": f"FreeLB halfway cavity (BGK<rhoU> + bounceback normal / movingwall managers), "
                             f"{n}^3 cells, {steps} timed steps after {warmup}, OpenMP {threads} threads, "
                             f"built with the benchmarks' flags; {float(fields[7]):.1f} s"}
            full = _fullway_reference(n, 20, threads)
            if full is not None:
                rec["fullway_value"] = full
                rec["fullway_sample"] = (f"FreeLB benchmarks/cavity3d (fullway walls), {n}^3 fluid cells, "
                                         "20 timed steps after 10 warm-up")
            return rec
    from oracle import oracle as O
    case, p = O.cavity_case("hw", n, n, n)
    case.L.nthreads = threads
    t0 = time.perf_counter()
    case.run(p, 20)
    dt = time.perf_counter() - t0
    return {"value": n ** 3 * 20 / dt / 1e6, "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"C oracle, halfway cavity {n}^3, 20 steps, OpenMP {threads} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--storages", default="f32,f16s,bf16s,f16c")
    ap.add_argument("--no-probe", action="store_true", help="skip the device copy-bandwidth probe")
    ap.add_argument("--precision", action="store_true",
                    help="run the 256^3 x 2000-step FP32/FP16/BF16 vs FP64 error study instead")
    ap.add_argument("--precision-steps", type=int, default=2000)
    ap.add_argument("--collisions", default="bgk,mrt_les", help="collisions of the precision study")
    ap.add_argument("--no-batch", action="store_true", help="skip the 512-block batched run")
    ap.add_argument("--config4", action="store_true",
                    help="512^3 MRT+LES cavity (Re 5000, lid 0.05), 5000 steps, f32/f16s/f16c vs FP64")
    ap.add_argument("--no-cavity", action="store_true", help="with --config4: only the batched adjoint")
    ap.add_argument("--mrt-les", action="store_true",
                    help="time the 512^3 MRT+LES cavity (Re 5000) for --steps steps per storage kind instead")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint runs")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the adjoint workloads (channel with and without recompute, 64-box batch), "
                         "e.g. under a profiler")
    ap.add_argument("--no-refined", action="store_true", help="skip the refined-cavity run")
    ap.add_argument("--no-coupled", action="store_true", help="skip the coupled flow + temperature run")
    ap.add_argument("--batch-eager", action="store_true",
                    help="time the batch with launches issued one by one (no hipGraph), e.g. under a profiler")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path here "
              "(each GPU would run its own share of blocks, or its own slab of one domain "
              "with a face exchange per step).", file=sys.stderr)
        sys.exit(2)

    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_bench()}), flush=True)
        print(json.dumps({"adjoint_no_recompute": adjoint_bench(K=0)}), flush=True)
        print(json.dumps({"adjoint_batch": adjoint_batch_bench()}), flush=True)
        print(json.dumps({"adjoint_batch_no_recompute": adjoint_batch_bench(K=0)}), flush=True)
        return
    if args.mrt_les:
        st = tuple(x for x in args.storages.split(",") if x != "bf16s")
        print(json.dumps({"mrt_les_512": mrt_les_speed(HipTimer(), storages=st, steps=args.steps,
                                                       warmup=args.warmup)}), flush=True)
        return
    if args.config4:
        # Re = u N / nu = 5000 with u = 0.05, N = 512: tau0 = 0.5 + 3 nu = 0.51536
        if not args.no_cavity:
            res = precision_study(n=512, steps=5000, collisions=("mrt_les",), storages=("f32", "f16s", "f16c"),
                                  u0_amp=0.0, u_lid=0.05, tau=0.5 + 3 * 0.05 * 512 / 5000)
            print(json.dumps({"config4_cavity": res}), flush=True)
        print(json.dumps({"config4_adjoint_batch": adjoint_batch_bench(fp64_check=True)}), flush=True)
        return
    if args.precision:
        st = tuple(s for s in args.storages.split(",") if s != "f64")
        res = precision_study(steps=args.precision_steps, collisions=tuple(args.collisions.split(",")),
                              storages=st)
        print(json.dumps({"precision_study": res}), flush=True)
        return
    timer = HipTimer()
    runs = {}
    for s in args.storages.split(","):
        runs[s] = time_storage(s, args.steps, args.warmup, timer)
        print(f"# {s}: {runs[s]['mlups']:.1f} MLUPS, {runs[s]['ms_per_step']:.4f} ms/step, "
              f"{runs[s]['achieved_gbs']:.0f} GB/s ({100 * runs[s]['frac']:.1f}% of HBM peak)",
              file=sys.stderr, flush=True)
    f32 = runs["f32"]
    line = {
        # the headline metric: FP32 MLUPS of the 256^3 cavity (value), with the
        # FP16-shift rate, the HBM roofline fraction and the adjoint MLUPS beside it
        "metric": "MLUPS D3Q19 256\u00b3 FP32 vs FP16-shift, % HBM roofline; adjoint MLUPS",
        "value": round(f32["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(f32["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (cavity flags, uniform equilibrium start)",
        "config": {"workload": "cavity D3Q19 BGK 256^3, halfway walls + moving lid, tau 0.53",
                   "cells": N_EDGE ** 3, "lattice": "D3Q19", "collision": "BGK",
                   "storage": "f32", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(f32["achieved_gbs"], 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(f32["frac"], 4),
                     "traffic": profiled_traffic("f32")},
    }
    # FreeLB's Timer::Print_MainLoopPerformance counts the void ghost layer too
    # (utils/timer.h:82-100, block_geometry3d.hh:448-460): the same run in that convention
    line["mlups_with_ghost_layer"] = round(f32["mlups"] * ((N_EDGE + 2) / N_EDGE) ** 3, 1)
    for s, r in runs.items():
        if s == "f32":
            continue
        line[f"{s}_mlups"] = round(r["mlups"], 1)
        line[f"{s}_speedup_vs_f32"] = round(r["mlups"] / f32["mlups"], 3)
        line[f"{s}_roofline_frac"] = round(r["frac"], 4)
    if not all(r["finite"] for r in runs.values()):
        line["warning"] = "non-finite mass after the timed run"
    if not args.no_probe:
        from freelb_amd import abi
        # the best 2 GiB -> 2 GiB copy over the probe shapes (16-byte accesses
        # in flight per lane, plain or non-temporal, one visit per tile)
        gbs = C.c_double()
        best, shape = 0.0, ""
        for unroll in (1, 2, 4):
            for nt in (0, 1):
                abi.check(abi.lib().flb_bandwidth_probe_ex(0, unroll, nt, 0, C.c_size_t(2 << 30), 10,
                                                           C.byref(gbs)))
                if gbs.value > best:
                    best, shape = gbs.value, f"copy unroll {unroll}, {'non-temporal' if nt else 'plain'}"
        line["device_copy_gbs"] = round(best, 1)
        line["device_copy_shape"] = shape
        line["roofline"]["frac_of_copy"] = round(f32["achieved_gbs"] / best, 4)
    if not args.no_batch:
        line["batch"] = batch_bench(timer, eager=args.batch_eager, graph_steps=50 if args.batch_eager else 500)
        print(f"# batch: {line['batch']}", file=sys.stderr, flush=True)
    if not args.no_refined:
        line["refined"] = refined_bench(timer)
        print(f"# refined: {line['refined']}", file=sys.stderr, flush=True)
    if not args.no_coupled:
        line["coupled"] = coupled_bench(timer)
        print(f"# coupled: {line['coupled']}", file=sys.stderr, flush=True)
    if not args.no_adjoint:
        line["adjoint"] = adjoint_bench()
        print(f"# adjoint: {line['adjoint']}", file=sys.stderr, flush=True)
        # the same channel without recompute: T moment records, 22.6 GB of HBM
        line["adjoint_no_recompute"] = adjoint_bench(K=0)
        print(f"# adjoint, no recompute: {line['adjoint_no_recompute']}", file=sys.stderr, flush=True)
        line["adjoint_batch"] = adjoint_batch_bench()
        print(f"# adjoint batch: {line['adjoint_batch']}", file=sys.stderr, flush=True)
        # the same batch without recompute: T moment records per box, 92 GB of HBM
        line["adjoint_batch_no_recompute"] = adjoint_batch_bench(K=0)
        print(f"# adjoint batch, no recompute: {line['adjoint_batch_no_recompute']}", file=sys.stderr, flush=True)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
