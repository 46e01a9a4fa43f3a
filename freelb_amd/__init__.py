"""MI355X (gfx950) engine for FreeLB's per-block lattice update.

The hot path -- the fused AA-pattern collide-and-stream step with halfway
bounce-back -- runs in hand-written HIP kernels behind the C ABI of
include/flb_abi.h (libflb_hip.so, built in-tree). This package holds the
kernels' sources (csrc/), the build recipe and a thin host-side mirror of
FreeLB's BlockLattice interface.
"""
from .abi import FLBError  # noqa: F401
from .adjoint import Adjoint, AdjointBatch  # noqa: F401
from .lattice import (BlockHalo, BlockLatManagerCoupling, BlockLattice, BlockLatticeManager, HalfwayBB,  # noqa: F401
                      Key_TypePair, Refinement, RefinementManager, weights, write_vtm)

__all__ = ["Adjoint", "AdjointBatch", "BlockHalo", "BlockLatManagerCoupling", "BlockLattice", "BlockLatticeManager", "HalfwayBB", "Key_TypePair", "Refinement", "RefinementManager", "weights", "write_vtm", "FLBError"]
