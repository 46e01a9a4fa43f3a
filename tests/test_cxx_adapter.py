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
