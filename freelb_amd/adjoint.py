"""Differentiable D3Q19 update on the GPU: forward BGK with body force and
Brinkman drag, hand-written adjoint, checkpointed recompute (the
flb_adjoint_* functions of include/flb_abi.h)."""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import abi


class Adjoint:
    """A box of nx*ny*nz cells (FreeLB index order) with forward and adjoint runs.

    The step is BGK with a constant body-force density `force` (FreeLB's
    BGKForce / ConstForce) plus a Brinkman drag -alpha*s*rho*u per cell; cells
    marked solid and cells beyond a non-periodic axis are halfway
    bounce-back walls. The objective is the time average over steps 1..T of
    the x-momentum summed over the plane x = outlet_x."""

    def __init__(self, dims: Sequence[int], omega: float, force=(0.0, 0.0, 0.0),
                 brinkman_alpha: float = 0.0, outlet_x: int | None = None,
                 periodic=(False, False, False), storage: str = "f32", stream=None):
        L = abi.lib()
        d = abi.AdjointDesc()
        d.nx, d.ny, d.nz = (int(v) for v in dims)
        for k in range(3):
            d.periodic[k] = int(bool(periodic[k]))
        d.storage = abi.STORAGE[storage]
        d.omega = float(omega)
        for k in range(3):
            d.force[k] = float(force[k])
        d.brinkman_alpha = float(brinkman_alpha)
        d.outlet_x = int(d.nx - 1 if outlet_x is None else outlet_x)
        d.stream = stream
        self.dims = (d.nx, d.ny, d.nz)
        self.N = d.nx * d.ny * d.nz
        self.storage = storage
        h = C.c_void_p()
        abi.check(L.flb_adjoint_create(C.byref(d), C.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            abi.lib().flb_adjoint_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_solid(self, solid: np.ndarray) -> None:
        m = np.ascontiguousarray(np.asarray(solid).reshape(-1) != 0, dtype=np.uint8)
        assert m.size == self.N
        abi.check(abi.lib().flb_adjoint_set_solid(self._h, abi.ptr(m, C.c_uint8)))

    def set_fraction(self, s: np.ndarray) -> None:
        a = np.ascontiguousarray(np.asarray(s, dtype=np.float64).reshape(-1))
        assert a.size == self.N
        abi.check(abi.lib().flb_adjoint_set_fraction(self._h, abi.ptr(a, C.c_double)))

    def set_fraction_at(self, cell: int, value: float) -> None:
        abi.check(abi.lib().flb_adjoint_set_fraction_at(self._h, int(cell), float(value)))

    def init_equilibrium(self, rho: np.ndarray, u: np.ndarray) -> None:
        rhou = np.ascontiguousarray(np.concatenate(
            [np.asarray(rho, dtype=np.float64).reshape(1, -1),
             np.asarray(u, dtype=np.float64).reshape(3, -1)], axis=0))
        abi.check(abi.lib().flb_adjoint_init_equilibrium(self._h, abi.ptr(rhou, C.c_double)))

    def forward(self, steps: int, checkpoint_every: int = -1) -> float:
        """Runs `steps` steps from f^0 and returns the objective J.

        checkpoint_every k > 0 keeps every k-th state; the backward pass then
        recomputes each segment of k states from its checkpoint (k near
        sqrt(steps) balances memory and recompute). 0 records every state's
        moments (rho - 1 and u, 4 values per cell) and recomputes nothing.
        -1 records nothing: objective only, no backward pass."""
        J = C.c_double()
        abi.check(abi.lib().flb_adjoint_forward(self._h, int(steps), int(checkpoint_every), C.byref(J)))
        return J.value

    def backward(self) -> np.ndarray:
        g = np.empty(self.N, dtype=np.float64)
        abi.check(abi.lib().flb_adjoint_backward(self._h, abi.ptr(g, C.c_double)))
        return g

    def state(self) -> np.ndarray:
        out = np.empty((19, self.N), dtype=np.float64)
        abi.check(abi.lib().flb_adjoint_state(self._h, abi.ptr(out, C.c_double)))
        return out

    def timing(self):
        f, b = C.c_double(), C.c_double()
        abi.check(abi.lib().flb_adjoint_timing(self._h, C.byref(f), C.byref(b)))
        return f.value, b.value


class AdjointBatch:
    """Several Adjoint boxes (any sizes, one storage kind) stepped together,
    one launch per time step over all of their cells. Each member keeps its
    own fields; after forward() every member also holds its own objective,
    final state and recorded run."""

    def __init__(self, members: Sequence[Adjoint], stream=None):
        L = abi.lib()
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m._h for m in self.members])
        h = C.c_void_p()
        abi.check(L.flb_adjoint_batch_create(arr, len(self.members), stream, C.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            abi.lib().flb_adjoint_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def forward(self, steps: int, checkpoint_every: int = -1) -> np.ndarray:
        """Every member's objective after `steps` steps (checkpoint_every as
        for Adjoint.forward)."""
        J = np.empty(len(self.members), dtype=np.float64)
        abi.check(abi.lib().flb_adjoint_batch_forward(self._h, int(steps), int(checkpoint_every),
                                                      abi.ptr(J, C.c_double)))
        return J

    def backward(self) -> list:
        g = np.empty(sum(m.N for m in self.members), dtype=np.float64)
        abi.check(abi.lib().flb_adjoint_batch_backward(self._h, abi.ptr(g, C.c_double)))
        out, off = [], 0
        for m in self.members:
            out.append(g[off:off + m.N].copy())
            off += m.N
        return out

    def timing(self):
        f, b = C.c_double(), C.c_double()
        abi.check(abi.lib().flb_adjoint_batch_timing(self._h, C.byref(f), C.byref(b)))
        return f.value, b.value
