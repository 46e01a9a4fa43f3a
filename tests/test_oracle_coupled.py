"""The oracle's coupled flow + scalar lattices (CPU).

FreeLB's own coupled run (tests/golden/natconv_*.npz, written by
oracle/_ref/natconv_ref from FreeLB's headers) must be reproduced bit for bit:
flow and temperature populations, RHO / VELOCITY / TEMP and SCALARFORCE.

Closed forms pin what the fixtures do not isolate:
  * D2Q5 / D3Q7 diffusion: a Fourier mode at rest decays with the lattice's
    diffusivity D = (tau - 1/2) cs_eff^2, cs_eff^2 = sum_i w_i c_ix^2 = 1/3
    (D2Q5) and 1/4 (D3Q7: FreeLB's weights 1/4, 1/8 with the cs2 = 1/3
    equilibrium);
  * advection: in a uniform VELOCITY field the scalar's centre moves with
    sum_i c_i feq_i / C = u (D2Q5) or 3/4 u (D3Q7) once the first moment
    has relaxed;
  * the zeroth moment of the equilibrium: sum_i feq_i = C for D2Q5, but
    C (1 - 3/8 |u|^2) for FreeLB's D3Q7, so a uniform scalar in a uniform
    flow decays as C_t = C_0 (1 - omega 3/8 |u|^2)^t on D3Q7 and stays put
    on D2Q5;
  * scalar conservation: adiabatic (normal bounce-back) walls keep sum C
    (D2Q5 in a random flow, D3Q7 at rest);
  * buoyancy with Guo forcing: in a periodic box of uniform scalar the
    momentum sum_i c_i f_i of every cell grows by exactly
    F = gbeta (C - C0) per step, along the last axis only.
"""
import os

import numpy as np
import pytest

from natconv_case import ALL, AABB, BB, VOID, fixtures, load, oracle_run
from oracle import oracle as O


@pytest.mark.parametrize("path", fixtures(), ids=os.path.basename)
def test_oracle_reproduces_freelb_coupled_run(path):
    fx = load(path)
    pf, pt, fields, T, sf = oracle_run(fx)
    assert np.array_equal(pf.view(np.uint32), fx["flow"].view(np.uint32))
    assert np.array_equal(pt.view(np.uint32), fx["temp"].view(np.uint32))
    f32 = fields.astype(np.float32)
    assert np.array_equal(f32[0], fx["rho"])
    assert np.array_equal(f32[1:], fx["u"])
    assert np.array_equal(T.astype(np.float32), fx["T"])
    assert np.array_equal(sf.astype(np.float32), fx["sforce"])


def _periodic_scalar(q, dims, omega, C0, u=None):
    """A fully periodic scalar lattice (one flag, SCALAR_BGK) at equilibrium of C0 at rest."""
    N = int(np.prod(dims))
    d = 3 if q == 7 else 2
    flags = np.full(N, AABB, np.uint8)
    uf = np.zeros(d * N) if u is None else np.repeat(np.asarray(u, float)[:, None], N, axis=1).reshape(-1)
    lat = O.Lattice(q, dims, flags, [(AABB, O.DYN_SCALAR_BGK)], omega=omega, rho_field=np.asarray(C0, float).copy(),
                    u_field=uf, overlap=0)
    w = np.array([1 / 3] + [1 / 6] * 4) if q == 5 else np.array([1 / 4] + [1 / 8] * 6)
    pops = np.ascontiguousarray(w[:, None] * np.asarray(C0, float)[None, :])
    return lat, pops


@pytest.mark.parametrize("q,cs2", [(5, 1.0 / 3.0), (7, 1.0 / 4.0)])
def test_scalar_diffusion_rate(q, cs2):
    tau = 0.8
    n = 64
    dims = (n, 4, 4) if q == 7 else (n, 4, 1)
    x = np.arange(n)
    k = 2 * np.pi / n
    C0 = np.broadcast_to(1.0 + 0.01 * np.cos(k * x), dims[::-1]).reshape(-1).copy()
    lat, pops = _periodic_scalar(q, dims, 1.0 / tau, C0)
    steps = 400
    lat.run(pops, steps)
    C = pops.sum(axis=0).reshape(dims[::-1])[0, 0]
    amp = 2.0 / n * np.abs(np.sum((C - 1.0) * np.cos(k * x)))
    D = (tau - 0.5) * cs2
    # second order in k: the discrete decay agrees with exp(-D k^2 t) to ~(k dx)^2
    assert amp == pytest.approx(0.01 * np.exp(-D * k * k * steps), rel=2e-3)
    # the other lattice's diffusivity is excluded by a wide margin
    other = (tau - 0.5) * (1.0 / 4.0 if cs2 > 0.3 else 1.0 / 3.0)
    assert abs(amp - 0.01 * np.exp(-other * k * k * steps)) > 20 * abs(amp - 0.01 * np.exp(-D * k * k * steps))


@pytest.mark.parametrize("q,speed", [(5, 1.0), (7, 0.75)])
def test_scalar_advection_speed(q, speed):
    n = 96
    dims = (n, n, 1) if q == 5 else (n, 24, 24)
    d = 3 if q == 7 else 2
    u = np.array([0.04, 0.0, 0.0][:d])
    grid = np.meshgrid(*[np.arange(m) for m in dims[::-1]], indexing="ij")
    zz, yy, xx = grid if len(grid) == 3 else (None, *grid)
    r2 = (xx - 30.0) ** 2 + (yy - dims[1] / 2) ** 2 + ((zz - dims[2] / 2) ** 2 if zz is not None else 0)
    C0 = (0.5 * np.exp(-r2 / 18.0)).reshape(-1)
    lat, pops = _periodic_scalar(q, dims, 1.0 / 0.7, C0, u=u)
    xs = np.broadcast_to(np.arange(n, dtype=float), dims[::-1]).reshape(-1)

    def centre(p):
        C = p.sum(axis=0)
        return np.sum(C * xs) / np.sum(C)

    lat.run(pops, 100)
    c1 = centre(pops)
    lat.run(pops, 200)
    c2 = centre(pops)
    # D3Q7's equilibrium loses O(u^2) of the scalar per step, which shifts the
    # ratio of first to zeroth moment at that order
    assert (c2 - c1) / 200 == pytest.approx(speed * u[0], rel=2e-4 if q == 5 else 2e-3)


@pytest.mark.parametrize("q,loss", [(5, 0.0), (7, 3.0 / 8.0)])
def test_scalar_equilibrium_zeroth_moment(q, loss):
    dims = (6, 5, 4) if q == 7 else (6, 5, 1)
    N = int(np.prod(dims))
    u = np.array([0.05, -0.03, 0.02][: 3 if q == 7 else 2])
    omega = 1.0 / 0.75
    lat, pops = _periodic_scalar(q, dims, omega, np.full(N, 1.25), u=u)
    steps = 40
    lat.run(pops, steps)
    expect = 1.25 * (1.0 - omega * loss * float(u @ u)) ** steps
    assert np.allclose(pops.sum(axis=0), expect, rtol=1e-13)


@pytest.mark.parametrize("q", [5, 7])
def test_scalar_conserved_by_adiabatic_walls(q):
    from oracle.oracle import lib  # noqa: F401  (builds the library)
    d = 3 if q == 7 else 2
    Ni = 12
    nz = Ni + 2 if d == 3 else 1
    dims = (Ni + 2, Ni + 2, nz)
    flags = np.zeros(int(np.prod(dims)), np.uint8)
    O.lib().flbo_cavity_flags(Ni, Ni, Ni if d == 3 else 0, flags.ctypes.data_as(O.C.POINTER(O.C.c_uint8)))
    flags = np.where(flags == O.BBMW, O.BB, flags).astype(np.uint8)
    N = flags.size
    rng = np.random.default_rng(3)
    C0 = 1.0 + 0.2 * rng.random(N)
    # a random flow on D2Q5; none on D3Q7, whose equilibrium itself loses
    # 3/8 |u|^2 of the scalar (see test_scalar_equilibrium_zeroth_moment)
    uf = 0.03 * (rng.random(d * N) - 0.5) if q == 5 else np.zeros(d * N)
    lat = O.Lattice(q, dims, flags, [(AABB | BB, O.DYN_SCALAR_BGK)], [(BB, VOID, O.HWBC_NORMAL)],
                    omega=1.0 / 0.9, rho_field=C0, u_field=uf)
    w = np.array([1 / 3] + [1 / 6] * 4) if q == 5 else np.array([1 / 4] + [1 / 8] * 6)
    pops = np.ascontiguousarray(w[:, None] * C0[None, :])
    live = flags != VOID
    m0 = pops[:, live].sum()
    lat.run(pops, 150)
    assert pops[:, live].sum() == pytest.approx(m0, rel=1e-13)


@pytest.mark.parametrize("qf,qt", [(9, 5), (19, 7)])
def test_buoyancy_momentum_grows_by_force(qf, qt):
    d = 3 if qf == 19 else 2
    dims = (8, 6, 5) if d == 3 else (8, 6, 1)
    N = int(np.prod(dims))
    flags = np.full(N, AABB, np.uint8)
    flow = O.Lattice(qf, dims, flags, [(ALL, O.DYN_BGK_SCALAR_FORCE)], omega=1.0 / 0.6, rho_field=np.ones(N),
                     u_field=np.zeros(d * N), overlap=0)
    T = np.full(N, 1.5)
    # the scalar held by its field (useFieldrhoU), so the buoyancy stays uniform
    temp = O.Lattice(qt, dims, flags, [(AABB, O.DYN_SCALAR_BGK_FIELD)], omega=1.0 / 0.7, rho_field=T.copy(),
                     overlap=0)
    from freelb_amd.lattice import weights
    pf = np.ascontiguousarray(np.repeat(weights(qf)[:, None], N, axis=1))
    pt = np.ascontiguousarray(weights(qt)[:, None] * T[None, :])
    gbeta, t0, steps = 1e-4, 1.0, 60
    O.run_coupled(flow, pf, [temp], [pt], [(0, ALL, gbeta, t0)], steps)
    cvec = _velocities(qf)
    mom = cvec.T @ pf  # [3][N]
    F = gbeta * (1.5 - t0)
    assert np.allclose(mom[d - 1], steps * F, rtol=1e-10, atol=1e-16)
    for a in range(d - 1):
        assert np.abs(mom[a]).max() < 1e-15
    assert np.allclose(flow.buffers["sforce"], F, rtol=1e-14)


def _velocities(q):
    if q == 19:
        return np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0],
                         [-1, -1, 0], [1, 0, 1], [-1, 0, -1], [0, 1, 1], [0, -1, -1], [1, -1, 0], [-1, 1, 0],
                         [1, 0, -1], [-1, 0, 1], [0, 1, -1], [0, -1, 1]], float)
    return np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [1, 1, 0], [-1, -1, 0], [1, -1, 0],
                     [-1, 1, 0]], float)
