"""The oracle's D3Q19 MRT and MRT + Smagorinsky collisions pinned by closed forms.

FreeLB has no D3Q19 MRT (its MRT_Feq_RhoU is written out for D2Q9 only,
lbm/collisionMRT.h:55-190), so nothing of the reference can be run against
these two collisions. They are pinned here instead by what the mathematics
fixes, built in the test from the published definitions rather than from the
oracle's tables:

* each of the 19 moments of d'Humieres, Ginzburg, Krafczyk, Lallemand & Luo
  (2002, Phil. Trans. R. Soc. A 360, 437-451) relaxes toward its equilibrium
  at its own rate: m'_j - m_eq_j = (1 - s_j) (m_j - m_eq_j), row by row, for
  the paper's default rates and for a seeded set of distinct rates (a rate
  applied to the wrong row fails it);
* in MRT + LES the five stress rows take the Smagorinsky omega that FreeLB's
  SmagorinskyBGK (lbm/collisionLES.h:29-67) computes for the same cell, and
  the others keep their rates; with every rate on the Smagorinsky omega the
  collision is SmagorinskyBGK, which the FreeLB-written smag fixture pins;
* a transverse shear wave decays at nu k^2 with nu = (1/omega - 1/2) / 3.

Single-cell collisions come from one oracle step on a small box of cells
with no walls: the step is collision then FreeLB's periodic-shift streaming
(a shift of each direction's whole array by Delta_i = c_i . (1, nx, nx ny)),
which the test undoes with a roll per direction.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# D3Q19 velocities in FreeLB's order (lbm/lattice_set.h:263-266)
C19 = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                [1, 1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, -1], [0, 1, 1], [0, -1, -1],
                [1, -1, 0], [-1, 1, 0], [1, 0, -1], [-1, 0, 1], [0, 1, -1], [0, -1, 1]])
W19 = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)

# The moment order of the rate array the library takes (flb_set_mrt_rates).
NAMES = ["rho", "e", "eps", "jx", "qx", "jy", "qy", "jz", "qz", "3pxx", "3pixx", "pww", "piww",
         "pxy", "pyz", "pxz", "mx", "my", "mz"]
CONSERVED = {"rho", "jx", "jy", "jz"}
STRESS = {"3pxx", "pww", "pxy", "pyz", "pxz"}  # relax with the shear-viscosity rate


def dhumieres_rows():
    """The moment polynomials of d'Humieres et al. (2002), eq. (2.4), by name."""
    x, y, z = (C19[:, a].astype(float) for a in range(3))
    c2 = x * x + y * y + z * z
    rows = {
        "rho": np.ones(19), "e": 19 * c2 - 30, "eps": (21 * c2 * c2 - 53 * c2 + 24) / 2,
        "jx": x, "qx": (5 * c2 - 9) * x, "jy": y, "qy": (5 * c2 - 9) * y, "jz": z, "qz": (5 * c2 - 9) * z,
        "3pxx": 3 * x * x - c2, "3pixx": (3 * c2 - 5) * (3 * x * x - c2),
        "pww": y * y - z * z, "piww": (3 * c2 - 5) * (y * y - z * z),
        "pxy": x * y, "pyz": y * z, "pxz": x * z,
        "mx": (y * y - z * z) * x, "my": (z * z - x * x) * y, "mz": (x * x - y * y) * z,
    }
    return np.array([rows[n] for n in NAMES])


def paper_rates(omega):
    """The paper's rates by moment name: s_e 1.19, s_eps = s_pi 1.4, s_q 1.2,
    s_m 1.98, the stress moments at the viscous rate omega."""
    by_name = {"e": 1.19, "eps": 1.4, "3pixx": 1.4, "piww": 1.4, "qx": 1.2, "qy": 1.2, "qz": 1.2,
               "mx": 1.98, "my": 1.98, "mz": 1.98}
    return np.array([0.0 if n in CONSERVED else (omega if n in STRESS else by_name[n]) for n in NAMES])


def feq_textbook(rho, u):
    """Second-order equilibrium w_i rho (1 + 3 c.u + 9/2 (c.u)^2 - 3/2 u^2)."""
    cu = C19 @ u
    uu = (u * u).sum(0)
    return W19[:, None] * rho[None, :] * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * uu[None, :])


def moments_of(f):
    rho = f.sum(0)
    return rho, (C19.T @ f) / rho


def collide_cells(f, tasks, omega, rates=None, dtype=np.float64, rho_field=None):
    """Collide every column of f (19 x n, n a multiple of 4) once with the oracle."""
    n = f.shape[1]
    dims = _box_dims(n)
    rf = np.ones(n) if rho_field is None else rho_field
    lat = O.Lattice(19, dims, np.full(n, O.AABB, np.uint8), tasks, [], omega=omega, mrt_rates=rates,
                    rho_field=rf, u_field=np.zeros((3, n)))
    out = lat.run(np.ascontiguousarray(f.astype(dtype)), 1)
    # the step streamed out[i][id] <- post[i][id - delta_i]
    return np.array([np.roll(out[i], -_delta(dims)[i]) for i in range(19)]).astype(np.float64)


def _box_dims(n):
    """An m x 2 x 2 box: its N = 4m exceeds every |Delta_i| <= 3m + 1, so the
    oracle's streaming is an exact cyclic shift of the whole array."""
    assert n % 4 == 0 and n >= 8
    return (n // 4, 2, 2)


def _delta(dims):
    nx, ny, _ = dims
    return C19[:, 0] + C19[:, 1] * nx + C19[:, 2] * nx * ny


def random_cells(n, seed, amp_u=0.05, amp_neq=0.01):
    """Equilibrium of random rho, u plus a random non-equilibrium part in
    every moment (mass and momentum of the cell are then recomputed)."""
    rng = np.random.default_rng(seed)
    rho = 1 + 0.05 * rng.standard_normal(n)
    u = amp_u * rng.standard_normal((3, n))
    f = feq_textbook(rho, u) + amp_neq * W19[:, None] * rng.standard_normal((19, n))
    return f


def check_relaxation(f, post, rates, tol=1e-13):
    """m'_j - m_eq_j == (1 - s_j)(m_j - m_eq_j) for every row j (rates: per
    row, or per row and cell)."""
    M = dhumieres_rows()
    rho, u = moments_of(f)
    rho2, u2 = moments_of(post)
    np.testing.assert_allclose(rho2, rho, rtol=1e-14)
    np.testing.assert_allclose(u2, u, rtol=0, atol=1e-14)
    meq = M @ feq_textbook(rho, u)
    d0 = M @ f - meq
    d1 = M @ post - meq
    s = np.asarray(rates, dtype=float)
    if s.ndim == 1:
        s = s[:, None]
    scale = np.abs(d0).max(axis=1)
    for j, name in enumerate(NAMES):
        if name in CONSERVED:
            assert np.abs(d1[j]).max() < 1e-14, name
            continue
        assert scale[j] > 1e-4, f"{name}: the test state does not excite this moment"
        expect = (1 - s[j]) * d0[j]
        err = np.abs(d1[j] - expect).max() / scale[j]
        assert err < tol, f"moment {name} (row {j}): relaxation error {err:.3e}"


def test_rows_are_orthogonal_with_the_paper_norms():
    """The rows built here are mutually orthogonal, with the paper's squared norms."""
    M = dhumieres_rows()
    G = M @ M.T
    assert np.allclose(G - np.diag(np.diag(G)), 0)
    norms = dict(zip(NAMES, np.diag(G)))
    assert norms["rho"] == 19 and norms["e"] == 2394 and norms["eps"] == 252
    assert norms["jx"] == 10 and norms["qx"] == 40 and norms["3pxx"] == 36 and norms["3pixx"] == 72
    assert norms["pww"] == 12 and norms["piww"] == 24 and norms["pxy"] == 4 and norms["mx"] == 8


def test_mrt_default_rates_relax_each_moment():
    omega = 1 / 0.6
    f = random_cells(64, seed=5)
    post = collide_cells(f, [(O.AABB, O.DYN_MRT)], omega)
    check_relaxation(f, post, paper_rates(omega))


def test_mrt_distinct_rates_relax_each_moment():
    """Nineteen distinct seeded rates: a rate read from the wrong row moves
    that row's relaxation off (1 - s_j)."""
    omega = 1 / 0.6
    rng = np.random.default_rng(17)
    rates = rng.uniform(0.3, 1.95, 19)
    f = random_cells(64, seed=6)
    post = collide_cells(f, [(O.AABB, O.DYN_MRT)], omega, rates=rates)
    check_relaxation(f, post, rates)
    # NaN entries take omega (the API's marker for the shear rates)
    rates2 = rates.copy()
    rates2[[9, 11, 13, 14, 15]] = np.nan
    post2 = collide_cells(f, [(O.AABB, O.DYN_MRT)], omega, rates=rates2)
    check_relaxation(f, post2, np.where(np.isnan(rates2), omega, rates2))


def _smag_omega_of_bgk(f, post_bgk):
    """The per-cell omega of a BGK update post = f - omega (f - feq), by least squares."""
    rho, u = moments_of(f)
    neq = f - feq_textbook(rho, u)
    om = ((f - post_bgk) * neq).sum(0) / (neq * neq).sum(0)
    resid = np.abs(f - om[None, :] * neq - post_bgk).max()
    assert resid < 1e-15, f"SmagorinskyBGK is not a BGK update ({resid:.2e})"
    return om


@pytest.mark.parametrize("tau0", [0.5005, 0.53])
def test_smag_mrt_stress_rows_take_the_smagorinsky_omega(tau0):
    omega, cs = 1 / tau0, 0.17
    f = random_cells(64, seed=7, amp_neq=0.02)
    rf = 1 + 0.01 * np.random.default_rng(8).standard_normal(64)  # the RHO field it divides by
    post_bgk = collide_cells(f, [(O.AABB, O.DYN_SMAG_BGK, cs)], omega, rho_field=rf)
    om_s = _smag_omega_of_bgk(f, post_bgk)
    assert np.all(om_s < omega) and np.all(om_s > 0.5 * omega)  # tau_t > 0, and the state excites it
    post = collide_cells(f, [(O.AABB, O.DYN_SMAG_MRT, cs)], omega, rho_field=rf)
    s = np.repeat(paper_rates(omega)[:, None], 64, axis=1)
    for j, name in enumerate(NAMES):
        if name in STRESS:
            s[j] = om_s
    check_relaxation(f, post, s)


def test_smag_mrt_with_every_rate_les_is_smag_bgk_fp64():
    omega, cs = 1 / 0.5005, 0.17
    f = random_cells(64, seed=9, amp_neq=0.02)
    post_bgk = collide_cells(f, [(O.AABB, O.DYN_SMAG_BGK, cs)], omega)
    post = collide_cells(f, [(O.AABB, O.DYN_SMAG_MRT, cs)], omega, rates=np.full(19, np.nan))
    np.testing.assert_allclose(post, post_bgk, rtol=0, atol=1e-15)


def test_smag_mrt_with_every_rate_les_reproduces_freelb_smag_fixture():
    """The FreeLB-written D3Q19 Smagorinsky cavity (SmagorinskyBGK fluid,
    fullway walls, 40 steps): the oracle's MRT + LES with every rate on the
    LES omega lands on it to FP32 rounding, while plain BGK (no eddy
    viscosity) is far off it -- so the bound separates the two."""
    fx = np.load(os.path.join(GOLDEN, "cavity_d3q19_smag_8x7x6_40.npz"))
    case, p = O.cavity_case("smag", 8, 7, 6)
    tasks = [(O.AABB, O.DYN_SMAG_MRT, case.tasks[0][2])] + case.tasks[1:]
    lat = O.Lattice(19, (case.nx, case.ny, case.nz), case.flags, tasks, [], omega=case.omega,
                    rho_field=case.rho_field, u_field=case.u_field, mrt_rates=np.full(19, np.nan))
    lat.run(p, 40)
    fluid = case.flags == O.AABB
    err = np.abs(p[:, fluid] - fx["pops"][:, fluid]).max()
    case_bgk, pb = O.cavity_case("fw", 8, 7, 6)
    case_bgk.run(pb, 40)
    les_effect = np.abs(pb[:, fluid] - fx["pops"][:, fluid]).max()
    print(f"MRT+LES (all rates LES) vs FreeLB smag fixture: {err:.3e}; plain BGK vs it: {les_effect:.3e}")
    # 40 steps of FP32 rounding of O(0.1) populations: ~40 * 2 ulp(0.1)
    assert err < 40 * 2 * np.spacing(np.float32(0.1))
    assert les_effect > 20 * err


def _shear_wave_nu(n, comp, omega, tasks, rates=None, steps=(100, 400)):
    """Decay rate of u_comp = A sin(2 pi x / n) in an n x 2 x 2 box: the
    periodic-shift streaming moves x by c_x modulo n, so a field of x alone
    streams as on a periodic axis."""
    N = 4 * n
    x = np.arange(N) % n
    k = 2 * np.pi / n
    rho = np.ones(N)
    u = np.zeros((3, N))
    u[comp] = 1e-4 * np.sin(k * x)
    f = np.ascontiguousarray(feq_textbook(rho, u))
    lat = O.Lattice(19, (n, 2, 2), np.full(N, O.AABB, np.uint8), tasks, [], omega=omega, mrt_rates=rates)
    amps = []
    done = 0
    for t in steps:
        lat.run(f, t - done)
        done = t
        _, uu = moments_of(f)
        amps.append(2 / N * (uu[comp] * np.sin(k * x)).sum())
    return -np.log(amps[1] / amps[0]) / (k * k * (steps[1] - steps[0]))


@pytest.mark.parametrize("comp,row", [(1, "pxy"), (2, "pxz")])
def test_mrt_shear_wave_decays_at_the_bgk_viscosity(comp, row):
    """u_y(x) and u_z(x) shear waves load the pxy and pxz stress moments;
    their decay gives nu = (1/omega - 1/2)/3, to the O(k^2) error of the
    lattice at 64 cells per wavelength."""
    omega = 1 / 0.6
    nu = (1 / omega - 0.5) / 3
    nu_mrt = _shear_wave_nu(64, comp, omega, [(O.AABB, O.DYN_MRT)])
    assert abs(nu_mrt / nu - 1) < 0.01, (row, nu_mrt, nu)
    # the same wave with that stress row at 1.19 instead of omega decays visibly slower
    rates = paper_rates(omega)
    rates[NAMES.index(row)] = 1.19
    nu_off = _shear_wave_nu(64, comp, omega, [(O.AABB, O.DYN_MRT)], rates=rates)
    assert nu_off / nu > 1.5
