"""The C++ host adapter (include/flb/block_lattice.hpp): a FreeLB-style driver
compiles against it with the host compiler, and on the GPU it produces the
same rho/u as the oracle."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "freelb_amd", "examples", "cavity3d")


def test_adapter_compiles_with_host_compiler():
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "cavity3d")
        subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror",
                        "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "freelb_amd", "examples", "cavity3d.cpp"),
                        "-L" + os.path.join(ROOT, "freelb_amd"), "-lflb_hip", "-o", out], check=True)
        assert os.path.exists(out)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["hw", "fw", "force", "batch"])
def test_adapter_driver_matches_oracle(mode):
    from oracle import oracle as O
    n, steps = 20, 57
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "rhou.bin")
        r = subprocess.run([EXE, mode, str(n), str(steps), out], check=True, capture_output=True,
                           text=True, timeout=120)
        print(r.stdout, flush=True)
        rhou = np.fromfile(out, dtype=np.float32).reshape(4, -1)
    case, p = O.cavity_case("hw" if mode == "batch" else mode, n, n, n)
    case.run(p, steps)
    ref = case.rhou(p)
    sel = case.flags == O.AABB if mode == "fw" else case.flags != O.VOID
    assert np.array_equal(rhou[:, sel], ref[:, sel])


REF_EXE = os.path.join(ROOT, "freelb_amd", "examples", "cavref3d")


def test_refined_example_compiles_with_host_compiler():
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "cavref3d")
        subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror",
                        "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "freelb_amd", "examples", "cavref3d.cpp"),
                        "-L" + os.path.join(ROOT, "freelb_amd"), "-lflb_hip", "-o", out], check=True)
        assert os.path.exists(out)


@pytest.mark.gpu
def test_refined_example_matches_oracle():
    """The cavref2d-style loop through flb::RefinedPair (D3Q19) against the
    oracle's refined cavity, bit for bit on every cell the GPU updates."""
    from oracle import oracle as O
    from refine_case import compare_masks
    counts = 13
    rc = O.RefinedCavity("d3q19", 12, 12, 12, 4, 1, 2)
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "pops.bin")
        r = subprocess.run([REF_EXE, "12", str(rc.origin[0]), str(rc.m), str(counts), out], check=True,
                           capture_output=True, text=True, timeout=120)
        print(r.stdout, flush=True)
        raw = np.fromfile(out, dtype=np.float32)
    nc, nf = rc.popsC.size, rc.popsF.size
    pc, pf = raw[:nc].reshape(rc.popsC.shape), raw[nc:nc + nf].reshape(rc.popsF.shape)
    rc.run(counts)
    cm, fm = compare_masks(rc)
    assert np.array_equal(pc[:, cm], rc.popsC[:, cm])
    assert np.array_equal(pf[:, fm], rc.popsF[:, fm])


NC_EXE = os.path.join(ROOT, "freelb_amd", "examples", "natconv3d")


def test_coupled_example_compiles_with_host_compiler():
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "natconv3d")
        subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror",
                        "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "freelb_amd", "examples", "natconv3d.cpp"),
                        "-L" + os.path.join(ROOT, "freelb_amd"), "-lflb_hip", "-o", out], check=True)
        assert os.path.exists(out)


@pytest.mark.gpu
def test_coupled_example_matches_oracle():
    """The cazsblock2d-style coupled loop through flb::BlockLatManagerCoupling
    (D3Q19 flow + D3Q7 temperature, buoyancy, hot / cold / adiabatic walls)
    against the oracle's coupled run, bit for bit."""
    from freelb_amd.cases import COLD, HOT, heated_cavity_flags
    from freelb_amd.lattice import weights
    from oracle import oracle as O
    n, steps = 10, 25
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "pops.bin")
        r = subprocess.run([NC_EXE, str(n), str(steps), out], check=True, capture_output=True, text=True,
                           timeout=120)
        print(r.stdout, flush=True)
        raw = np.fromfile(out, dtype=np.float32)
    flags = heated_cavity_flags(n, 3)
    N = flags.size
    dims = (n + 2,) * 3
    ALL = 2 | 4 | HOT | COLD
    T0 = np.where(flags == HOT, 1.5, np.where(flags == COLD, 0.5, 1.0))
    flow = O.Lattice(19, dims, flags, [(ALL, O.DYN_BGK_SCALAR_FORCE)], [(4 | HOT | COLD, 1, O.HWBC_NORMAL)],
                     omega=1.0 / 0.6, rho_field=np.ones(N), u_field=np.zeros(3 * N))
    temp = O.Lattice(7, dims, flags, [(2 | 4, O.DYN_SCALAR_BGK), (HOT | COLD, O.DYN_SCALAR_BGK_FIELD)],
                     [(4, 1, O.HWBC_NORMAL), (HOT | COLD, 1, O.HWBC_ANTI)], omega=1.0 / 0.7, rho_field=T0)
    pf = np.ascontiguousarray(np.repeat(weights(19, np.float32)[:, None], N, axis=1))
    pt = np.ascontiguousarray((weights(7, np.float32)[:, None] * T0.astype(np.float32)[None, :]))
    O.run_coupled(flow, pf, [temp], [pt], [(0, ALL, 1e-4, 1.0)], steps)
    live = flags != 1
    assert np.array_equal(raw[:19 * N].reshape(19, N)[:, live], pf[:, live])
    assert np.array_equal(raw[19 * N:].reshape(7, N)[:, live], pt[:, live])
