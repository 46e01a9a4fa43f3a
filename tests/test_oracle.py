"""The oracle (C restatement of FreeLB's update) pinned to FreeLB itself.

Golden fixtures in tests/golden/ were written by FreeLB's own code (see
tests/golden/make_fixtures.py). The oracle must reproduce them bit for bit --
every population of every cell, void cells included, since the oracle also
restates FreeLB's periodic-shift streaming of the void layer.

Closed-form checks pin what the fixtures leave open (MRT, Smagorinsky, FP64):
equilibrium moments, rest-state fixed point, MRT == BGK when every rate is
omega, Smagorinsky with C_s = 0 == BGK, exact mass conservation in FP64.
"""
import glob
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests import shipped_case

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _small_fixtures():
    return sorted(p for p in glob.glob(os.path.join(GOLDEN, "cavity_d*_[0-9]*x*x*_[0-9]*.npz"))
                  if not p.endswith(("_midplane.npz", "_shipped.npz")))


def _case_of(path):
    base = os.path.basename(path)[len("cavity_"):-len(".npz")]
    lat, mode, dims, steps = base.split("_")
    ni, nj, nk = (int(v) for v in dims.split("x"))
    return lat, mode, ni, nj, (nk if lat == "d3q19" else 0), int(steps)


@pytest.mark.parametrize("path", _small_fixtures(), ids=lambda p: os.path.basename(p))
def test_oracle_reproduces_reference_bitwise(path):
    lat, mode, ni, nj, nk, steps = _case_of(path)
    fx = np.load(path)
    case, p = O.cavity_case(mode, ni, nj, nk)
    assert np.array_equal(case.flags, fx["flags"]), "cavity flag layout differs from FreeLB's"
    om, uw = O.convert_from_rt(np.float32(0.53), np.float32(0.01), np.float32(0.1))
    assert om == fx["omega"] and uw == fx["uw"]
    case.run(p, steps)
    assert np.array_equal(p, fx["pops"]), f"max |diff| {np.abs(p - fx['pops']).max()}"
    rh = case.rhou(p)
    nonvoid = case.flags != O.VOID
    assert np.array_equal(rh[:, nonvoid], fx["rhou"][:, nonvoid])


@pytest.mark.parametrize("mode", ["hw", "fw"])
def test_oracle_config0_100cubed_1000_steps(mode):
    """The 100^3 cavity (tau 0.53, lid 0.1) after 1000 steps: hash of every
    population of every cell equals FreeLB's."""
    meta = json.load(open(os.path.join(GOLDEN, "cavity_large.json")))[f"d3q19_{mode}_100x100x100_1000"]
    case, p = O.cavity_case(mode, 100, 100, 100)
    case.run(p, 1000)
    assert hashlib.sha256(p.tobytes()).hexdigest() == meta["sha256_pops_all_cells"]
    rh = case.rhou(p)
    nonvoid = case.flags != O.VOID
    assert hashlib.sha256(rh[:, nonvoid].tobytes()).hexdigest() == meta["sha256_rhou_nonvoid"]
    if mode == "hw":
        # FreeLB built as its benchmarks ship (contraction on): within the
        # relative bar on the sampled cells, and the recorded full-field
        # difference between its two builds within the bar as well
        e = shipped_case.errors(rh)
        print(e)
        assert e["rho_max_rel"] <= shipped_case.BAR and e["u_max_over_umax"] <= shipped_case.BAR
        full = e["full_field_shipped_vs_contraction_off"]
        assert full["rho_max_rel"] <= shipped_case.BAR and full["u_max_over_umax"] <= shipped_case.BAR


def test_cavity_flags_rule_matches_fixture_shapes():
    # non-cubic boxes: flag rule restated from the AABB / SetupBoundary logic
    for path in _small_fixtures():
        lat, mode, ni, nj, nk, steps = _case_of(path)
        assert np.array_equal(O.cavity_flags(ni, nj, nk), np.load(path)["flags"])


# ---- closed forms -----------------------------------------------------------

def _lattice(q):
    if q == 19:
        c = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                      [1, 1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, -1], [0, 1, 1], [0, -1, -1],
                      [1, -1, 0], [-1, 1, 0], [1, 0, -1], [-1, 0, 1], [0, 1, -1], [0, -1, 1]])
    else:
        c = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [1, 1, 0],
                      [-1, -1, 0], [1, -1, 0], [-1, 1, 0]])
    return c


def _box(q, n=6, tasks=None, bcs=None, rates=None):
    nk = n if q == 19 else 0
    flags = O.cavity_flags(n, n, nk)
    dims = (n + 2, n + 2, (n + 2) if q == 19 else 1)
    tasks = tasks or [(O.AABB | O.BB | O.BBMW, O.DYN_BGK)]
    bcs = bcs if bcs is not None else [(O.BB | O.BBMW, O.VOID, O.HWBC_NORMAL)]
    return O.Lattice(q, dims, flags, tasks, bcs, omega=1.0 / 0.6, mrt_rates=rates)


@pytest.mark.parametrize("q", [9, 19])
def test_equilibrium_moments_closed_form(q):
    lat = _box(q)
    rng = np.random.default_rng(1)
    d = 3 if q == 19 else 2
    rhou = np.zeros((d + 1, lat.N))
    rhou[0] = 1 + 0.05 * rng.standard_normal(lat.N)
    rhou[1:] = 0.05 * rng.standard_normal((d, lat.N))
    f = lat.init_feq(rhou)
    c = _lattice(q)[:, :d]
    rho = f.sum(0)
    j = c.T @ f
    np.testing.assert_allclose(rho, rhou[0], rtol=1e-13)
    np.testing.assert_allclose(j, rhou[0] * rhou[1:], rtol=1e-11, atol=1e-15)
    for a in range(d):
        for b in range(d):
            pab = (c[:, a] * c[:, b]) @ f
            expect = rhou[0] * rhou[1 + a] * rhou[1 + b] + (rhou[0] / 3.0 if a == b else 0.0)
            np.testing.assert_allclose(pab, expect, rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize("q", [9, 19])
def test_rest_state_is_fixed_point(q):
    lat = _box(q)
    w = O.weights(q, np.float64)
    p = np.ascontiguousarray(np.repeat(w[:, None], lat.N, axis=1))
    p0 = p.copy()
    lat.run(p, 17)
    nonvoid = lat.flags != O.VOID
    np.testing.assert_allclose(p[:, nonvoid], p0[:, nonvoid], rtol=0, atol=1e-15)


def _random_state(lat, seed=3, amp=0.02):
    rng = np.random.default_rng(seed)
    d = lat.d
    rhou = np.zeros((d + 1, lat.N))
    rhou[0] = 1 + amp * rng.standard_normal(lat.N)
    rhou[1:] = amp * rng.standard_normal((d, lat.N))
    return lat.init_feq(rhou)


@pytest.mark.parametrize("q", [19])
def test_mrt_with_all_rates_omega_equals_bgk(q):
    """D3Q19 MRT (projected equilibrium moments, orthogonal basis) reduces to
    BGK exactly when every rate is omega. FreeLB's D2Q9 MRT keeps O(u^3)
    terms in its equilibrium moments (collisionMRT.ur.h:96-104), so it does
    not reduce; the D2Q9 form is pinned by its golden fixture instead."""
    om = 1.0 / 0.6
    rates = np.full(q, np.nan)  # every rate takes omega
    mrt = _box(q, tasks=[(O.AABB | O.BB | O.BBMW, O.DYN_MRT)], rates=rates)
    bgk = _box(q)
    p1 = _random_state(bgk)
    p2 = p1.copy()
    bgk.run(p1, 9)
    mrt.run(p2, 9)
    nonvoid = bgk.flags != O.VOID
    np.testing.assert_allclose(p2[:, nonvoid], p1[:, nonvoid], rtol=0, atol=1e-14)
    assert om > 0


@pytest.mark.parametrize("q", [9, 19])
@pytest.mark.parametrize("dyn", ["mrt", "smag"])
def test_fp64_mass_conservation_closed_box(q, dyn):
    kind = O.DYN_MRT if dyn == "mrt" else O.DYN_SMAG_BGK
    lat = _box(q, n=7, tasks=[(O.AABB | O.BB | O.BBMW, kind, 0.17)])
    p = _random_state(lat)
    nonvoid = lat.flags != O.VOID
    m0 = p[:, nonvoid].sum()
    c = _lattice(q)
    lat.run(p, 25)
    assert abs(p[:, nonvoid].sum() - m0) < 1e-12 * m0
    assert np.isfinite(p).all()
    assert c.shape[0] == q


@pytest.mark.parametrize("q", [9, 19])
def test_smagorinsky_cs0_equals_bgk(q):
    smag = _box(q, tasks=[(O.AABB | O.BB | O.BBMW, O.DYN_SMAG_BGK, 0.0)])
    bgk = _box(q)
    p1 = _random_state(bgk)
    p2 = p1.copy()
    bgk.run(p1, 9)
    smag.run(p2, 9)
    nonvoid = bgk.flags != O.VOID
    # 1/(1/omega) is not omega bit for bit; the difference is one rounding
    np.testing.assert_allclose(p2[:, nonvoid], p1[:, nonvoid], rtol=1e-14, atol=1e-16)


def test_d3q19_mrt_basis_is_orthogonal():
    """The d'Humieres moment rows used by the oracle are mutually orthogonal."""
    c = _lattice(19)
    x, y, z = c[:, 0], c[:, 1], c[:, 2]
    c2 = x * x + y * y + z * z
    M = np.array([np.ones(19), 19 * c2 - 30, (21 * c2 * c2 - 53 * c2 + 24) / 2, x, (5 * c2 - 9) * x, y,
                  (5 * c2 - 9) * y, z, (5 * c2 - 9) * z, 3 * x * x - c2, (3 * c2 - 5) * (3 * x * x - c2),
                  y * y - z * z, (3 * c2 - 5) * (y * y - z * z), x * y, y * z, x * z,
                  (y * y - z * z) * x, (z * z - x * x) * y, (x * x - y * y) * z])
    G = M @ M.T
    assert np.allclose(G - np.diag(np.diag(G)), 0)


# ---- two-level refinement --------------------------------------------------------

REFINED_FIXTURES = sorted(glob.glob(os.path.join(GOLDEN, "refined_*.npz")))


@pytest.mark.parametrize("path", REFINED_FIXTURES, ids=os.path.basename)
def test_refined_cavity_reproduces_reference_bitwise(path):
    """FreeLB's refined cavity (examples/cavref2d/cavref2d.cpp's loop, several
    coarse blocks around one fine block) against the oracle's one coarse and
    one fine block: every population and field of every cell, bit for bit."""
    from refine_case import fixture_mismatches, load_fixture
    fx = load_fixture(path)
    rc = O.RefinedCavity("d2q9", fx["n"], fx["n"], 1, fx["cell"], fx["p0"], fx["p1"])
    rc.run(fx["counts"])
    bad, checked = fixture_mismatches(rc, fx)
    assert checked > 0 and bad == 0, f"{bad} of {checked} cells differ"


@pytest.mark.parametrize("lattice", ["d2q9", "d3q19"])
def test_refinement_keeps_rest_state(lattice):
    """No lid motion: the rest state f = w is a fixed point of both levels and
    of both conversions (to FP64 rounding)."""
    rc = O.RefinedCavity(lattice, 12, 12, 12, 4, 1, 2, u_phys=0.0, dtype=np.float64)
    w = O.weights(rc.q, np.float64)
    rc.run(7)
    cm = (rc.coarse.flags != O.VOID) & (rc.coarse.flags != O.COVERED)
    np.testing.assert_allclose(rc.popsC[:, cm], np.repeat(w[:, None], int(cm.sum()), axis=1), rtol=0, atol=1e-15)
    np.testing.assert_allclose(rc.popsF, np.repeat(w[:, None], rc.popsF.shape[1], axis=1), rtol=0, atol=1e-15)


def _linear(q, pts, a, b):
    return a[:, None] + b @ pts.T


@pytest.mark.parametrize("lattice", ["d2q9", "d3q19"])
def test_refinement_reproduces_linear_fields(lattice):
    """With rho = 1 and u = 0 in the fields feq = w, so both conversions are
    affine maps of the exchanged value; a population field linear in space
    must come out as the same linear field at the fine (coarse) cell centres,
    which fixes the stencil placement, the weights and which sub-cell each
    fine cell is (a swapped sub-cell or a transposed weight breaks it)."""
    rc = O.RefinedCavity(lattice, 12, 12, 12, 4, 1, 2, u_phys=0.0, dtype=np.float64)
    d, q = rc.d, rc.q
    rng = np.random.default_rng(5)
    a = rng.uniform(0.5, 1.0, q)
    b = rng.uniform(-0.01, 0.01, (q, d))
    w = O.weights(q, np.float64)

    def centres(dims, origin_shift, h):
        idx = np.indices(dims[::-1]).reshape(len(dims), -1)[::-1][:d].T  # (N, d), x first
        return origin_shift + (idx + 0.5) * h

    # coarse cell k spans [k, k + 1); fine cell f of the patch at origin o spans
    # o - 1 + (f + [0, 1]) / 2
    cpts = centres(rc.cdims, 0.0, 1.0)
    rc.popsC[:] = _linear(q, cpts, a, b)
    O.refine_interpolate(rc.coarse, rc.popsC, rc.fine, rc.popsF, rc.origin)
    fpts = centres(rc.fdims, np.array(rc.origin[:d], dtype=float) - 1.0, 0.5)
    want = _linear(q, fpts, a, b)
    want = want - 0.25 * rc.omega_c * (want - w[:, None])
    fn = rc.fdims
    ghost = np.zeros(fn[::-1], dtype=bool)
    if d == 3:
        ghost[...] = True
        ghost[2:-2, 2:-2, 2:-2] = False
    else:
        ghost[0] = True
        ghost[0, 2:-2, 2:-2] = False
    ghost = ghost.reshape(-1)
    np.testing.assert_allclose(rc.popsF[:, ghost], want[:, ghost], rtol=0, atol=1e-13)

    rc.popsF[:] = _linear(q, fpts, a, b)
    O.refine_average(rc.coarse, rc.popsC, rc.fine, rc.popsF, rc.origin)
    want = _linear(q, cpts, a, b)
    want = want + 0.5 * rc.omega_f * (want - w[:, None])
    o, m = rc.origin[0], rc.m
    ring = np.zeros(rc.cdims[::-1], dtype=bool)
    box = (slice(o, o + m),) * d
    inner = (slice(o + 1, o + m - 1),) * d
    if d == 3:
        ring[box] = True
        ring[inner] = False
    else:
        ring[(0,) + box] = True
        ring[(0,) + inner] = False
    ring = ring.reshape(-1)
    np.testing.assert_allclose(rc.popsC[:, ring], want[:, ring], rtol=0, atol=1e-13)
