"""Convergence residual (flb_enable_tolerance / flb_tolerance, FreeLB's
getToleranceU / getTolU(shift) and the RHO forms): on a cavity whose cells
write their moments to the fields (BGK<rhoU<CELL, true>>), the GPU fields
equal the oracle's bit for bit, and the residual equals the maximum
|field(t2) - field(t1)| taken from the oracle's fields at the two times."""
import numpy as np
import pytest

from gpu_case import gpu_lattice
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _case(q):
    n = (12, 10, 9) if q == 19 else (16, 13, 0)
    case, p = O.cavity_case("hwfield", *n)
    lat = O.Lattice(case.q, (case.nx, case.ny, case.nz), case.flags,
                    [(O.AABB, O.DYN_BGK_WRITE), (O.BB | O.BBMW, O.DYN_BGK_FIELD)], case.bcs, omega=case.omega,
                    rho_field=case.rho_field, u_field=case.u_field)
    return lat, p


def _residual(new, old, dims, shift):
    nx, ny, nz = dims
    a = np.abs(new.reshape(-1, nz, ny, nx).astype(np.float32) - old.reshape(-1, nz, ny, nx).astype(np.float32))
    zs = slice(shift, nz - shift) if nz > 1 else slice(0, 1)
    return float(a[:, zs, shift:ny - shift, shift:nx - shift].max())


@pytest.mark.parametrize("q", [9, 19])
@pytest.mark.parametrize("shift", [0, 1, 2])
def test_tolerance_matches_oracle_fields(q, shift):
    lat, p = _case(q)
    g = gpu_lattice(lat, "f32", p)
    try:
        lat.run(p, 3)
        g.step(3)
        f1 = lat.fields().copy()
        assert np.array_equal(g.download_fields(), f1)
        g.enable_tolerance()
        for k in (1, 4, 7):
            lat.run(p, k)
            g.step(k)
            f2 = lat.fields().copy()
            assert np.array_equal(g.download_fields(), f2)
            dims = (lat.nx, lat.ny, lat.nz)
            ru = g.tolerance("u", shift)
            rr = g.tolerance("rho", shift)
            assert ru == _residual(f2[1:], f1[1:], dims, shift)
            assert rr == _residual(f2[:1], f1[:1], dims, shift)
            assert ru > 0.0
            f1 = f2
        # nothing changed since the last call
        assert g.tolerance("u", shift) == 0.0
    finally:
        g.close()
