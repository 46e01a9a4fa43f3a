"""Builds libflb_hip.so (the gfx950 kernels and the C ABI) in-tree with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libflb_hip.so")
SOURCES = ["flb_inst_f32.hip", "flb_inst_f16s.hip", "flb_inst_bf16s.hip", "flb_inst_f64.hip", "flb_inst_f16c.hip",
           "flb_probe.hip", "flb_links.hip", "flb_abi.hip", "flb_batch.hip", "flb_halo.hip", "flb_adjoint.hip",
           "flb_refine.hip", "flb_tol.hip", "flb_vtk.hip"]
HEADERS = ["flb_cell.hpp", "flb_storage.hpp", "flb_internal.hpp", "flb_kernels.hpp", "flb_host.hpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FLB_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # FP32 results must round exactly as FreeLB's do: no FMA contraction and
    # IEEE division / square root.
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    # no packed FP32 instructions on this hardware pool
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


# Per-source flags. The 16-bit storage kinds have no FreeLB counterpart to
# match bit for bit, so their collision arithmetic may contract to FMAs and
# use the hardware reciprocal: their kernels are instruction-bound rather
# than memory-bound. Contraction stays within one source expression
# (-ffp-contract=on, as fast as =fast here), so the same cell update inlined
# into different kernels (the one-cell and two-cell steps, batches) rounds
# the same way and their results agree bit for bit.
SOURCE_FLAGS = {
    # The differentiable step has no FreeLB counterpart either (its FP64
    # model is the reference); its one division per cell stays IEEE.
    "flb_adjoint.hip": ["-ffp-contract=on"],
    "flb_inst_f16s.hip": ["-ffp-contract=on", "-fno-hip-fp32-correctly-rounded-divide-sqrt"],
    "flb_inst_bf16s.hip": ["-ffp-contract=on", "-fno-hip-fp32-correctly-rounded-divide-sqrt"],
    "flb_inst_f16c.hip": ["-ffp-contract=on", "-fno-hip-fp32-correctly-rounded-divide-sqrt"],
}


def build(verbose: bool = False, jobs: int = 0, extra_flags=(), lib: str = LIB,
          objdir: str = "") -> str:
    """Compile every source to an object in parallel, then link the library.
    `extra_flags` (e.g. -D switches), `lib` and `objdir` build a variant
    library beside the default one."""
    from concurrent.futures import ThreadPoolExecutor

    objdir = objdir or os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    compile_flags = [f for f in FLAGS if f != "-shared"]

    def compile_one(src):
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, *compile_flags, *SOURCE_FLAGS.get(src, []), *extra_flags, "-c", "-o", obj,
               os.path.join(CSRC, src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        return obj

    jobs = jobs or min(len(SOURCES), os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    # roctx: named ranges around the host calls (flbh::Range)
    link = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", lib, *objs, "-L/opt/rocm/lib",
            "-lrocprofiler-sdk-roctx"]
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.run(link, check=True)
    return lib


EXAMPLES = {"cavity3d": "cavity3d.cpp", "cavref3d": "cavref3d.cpp", "natconv3d": "natconv3d.cpp"}


def build_examples(verbose: bool = False) -> list:
    """C++ drivers over include/flb/block_lattice.hpp, built with the host compiler."""
    root = os.path.dirname(HERE)
    out = []
    for name, src in EXAMPLES.items():
        exe = os.path.join(HERE, "examples", name)
        cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-Wall", "-Wextra",
               "-I" + os.path.join(root, "include"), os.path.join(HERE, "examples", src),
               "-L" + HERE, "-lflb_hip", "-Wl,-rpath,$ORIGIN/..", "-o", exe]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        out.append(exe)
    return out


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    build_examples(verbose="-v" in sys.argv)
    print(LIB)
