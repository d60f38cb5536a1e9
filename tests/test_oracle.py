"""Pins for the float64 oracle: closed forms from prolix's own tests, finite
differences, invariants, CMAP interpolation properties, Philox known answers."""

from __future__ import annotations

import math
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import gb_oracle as O
from oracle import philox
from tests.helpers import protein

ROOT = Path(__file__).resolve().parents[1]


# ---------------------------------------------------------------- closed forms

def test_born_ion():
    """tests/physics/test_generalized_born.py:11-52: B = R - 0.09, E = -1/2 k tau q^2 / B."""
    B = O.born_radii(np.zeros((1, 3)), np.array([2.0]))
    assert B[0] == pytest.approx(1.91, rel=1e-14)
    e = O.gb_energy(np.zeros((1, 3)), np.array([1.0]), B, eps_in=1.0, eps_out=78.5)
    assert e == pytest.approx(-0.5 * 332.0637 * (1 - 1 / 78.5) / 1.91, rel=1e-9)


def test_bond_energy_and_force():
    """tests/physics/test_bonded.py:10-38: k = 100, dr = 0.1 -> E = 0.5, F = -/+10."""
    s = {"charges": np.zeros(2), "sigmas": np.ones(2), "epsilons": np.zeros(2), "radii": np.ones(2) * 1.5,
         "bonds": np.array([[0, 1]]), "bond_params": np.array([[1.0, 100.0]])}
    pos = np.array([[0.0, 0, 0], [1.1, 0, 0]])
    t = O.energy_terms(s, pos, implicit_solvent=False)
    assert t["bond"] == pytest.approx(0.5, rel=1e-12)
    f = O.forces(s, pos, implicit_solvent=False)
    assert f[1, 0] == pytest.approx(-10.0, abs=1e-9)
    assert f[0, 0] == pytest.approx(10.0, abs=1e-9)


def test_angle_energy():
    """tests/physics/test_bonded.py:41-59: 90 degrees vs theta0 = pi, k = 100."""
    pos = np.array([[1.0, 0, 0], [0.0, 0, 0], [0.0, 1, 0]])
    e = O.angle_energy(pos, np.array([[0, 1, 2]]), np.array([[math.pi, 100.0]]))
    assert e == pytest.approx(0.5 * 100 * (math.pi / 2) ** 2, rel=1e-4)


def test_lj_pair_value():
    """tests/physics/test_numerical_stability.py:9-26: r = 1.5, sigma = eps = 1 -> -0.3203."""
    pos = np.array([[0.0, 0, 0], [1.5, 0, 0]])
    S = np.ones((2, 2))
    e = O.lj_energy(pos, np.ones(2), np.ones(2), S, cutoff=2.0)
    assert abs(e + 0.3203) < 1e-4
    # overlapping atoms are masked, not NaN (same test, :28-36)
    e0 = O.lj_energy(np.zeros((2, 3)), np.ones(2), np.ones(2), S, cutoff=2.0)
    assert e0 == 0.0


def test_coulomb_pair_and_exclusion_scale():
    pos = np.array([[0.0, 0, 0], [2.0, 0, 0]])
    q = np.array([0.5, -0.4])
    assert O.coulomb_energy(pos, q, np.ones((2, 2))) == pytest.approx(332.0637 * -0.2 / 2.0, rel=1e-14)
    idx = np.array([[1, -1], [0, -1]])
    sc = np.array([[0.8333, 1.0], [0.8333, 1.0]])
    S = O.dense_scales(idx, sc, 2)
    assert O.coulomb_energy(pos, q, S) == pytest.approx(0.8333 * 332.0637 * -0.2 / 2.0, rel=1e-14)


def test_dihedral_cis_trans():
    cis = np.array([[1.0, 1, 0], [0.0, 0, 0], [0.0, 0, 1], [1.0, 1, 1]])
    trans = np.array([[1.0, 0, 0], [0.0, 0, 0], [0.0, 0, 1], [-1.0, 0, 1]])
    idx = np.array([[0, 1, 2, 3]])
    assert O.dihedral_angles(cis, idx)[0] == pytest.approx(0.0, abs=1e-7)
    assert abs(O.dihedral_angles(trans, idx)[0]) == pytest.approx(math.pi, abs=1e-7)


def test_pair_integral_limits_and_far_field_identity():
    # vanishes far away, zero when j's sphere sits inside i's offset sphere
    assert O.pair_integral(np.array(1e4), 1.5, 1.2) < 1e-12
    assert O.pair_integral(np.array(0.1), 1.5, 0.2) == 0.0
    # far field (r - s > rho): I = (x / (1 - x^2) - atanh x) / 2r, x = s / r
    rng = np.random.default_rng(0)
    for _ in range(200):
        s, rho = rng.uniform(0.8, 1.8), rng.uniform(1.1, 1.9)
        r = s + rho + rng.uniform(0.01, 40)
        x = s / r
        assert O.pair_integral(np.array(r), rho, s) == pytest.approx((x / (1 - x * x) - math.atanh(x)) / (2 * r),
                                                                     rel=1e-9, abs=1e-15)


@pytest.mark.parametrize("suffix", ["", "4"])
def test_far_field_polynomials_match_series(suffix):
    """The coefficients compiled into the pair kernels reproduce h and k: the
    fits on s <= r/2 and the degree-4 fits on s <= r/4."""
    txt = (ROOT / "prolix_amd/csrc/prx_obc_farfield.h").read_text()
    H = [float(v.rstrip("f")) for v in re.search(rf"kObcH{suffix}\[\d+\] = \{{(.*?)\}}", txt).group(1).split(",")]
    K = [float(v.rstrip("f")) for v in re.search(rf"kObcK{suffix}\[\d+\] = \{{(.*?)\}}", txt).group(1).split(",")]
    ymax = float(re.search(rf"PRX_OBC_FAR{suffix}_YMAX ([0-9.]+)f", txt).group(1))
    xmax = float(re.search(rf"PRX_OBC_FAR{suffix}_XMAX ([0-9.]+)f", txt).group(1))
    assert (len(H), len(K)) == ((7, 8) if suffix == "" else (5, 5))
    assert xmax * xmax == pytest.approx(ymax, rel=1e-7)
    y = np.linspace(0.0, ymax, 2001)
    h = sum(2 * k / (2 * k + 1) * y ** (k - 1) for k in range(1, 300))
    kk = sum(2 * k * (k + 1) / (2 * k + 1) * y ** (k - 1) for k in range(1, 300))
    assert np.max(np.abs(np.polyval(H[::-1], y) / h - 1)) < 2e-7
    assert np.max(np.abs(np.polyval(K[::-1], y) / kk - 1)) < 2e-7
    # and the polynomial form equals the oracle's integral and its derivative
    rng = np.random.default_rng(1)
    for _ in range(100):
        s, rho = rng.uniform(0.8, 1.8), rng.uniform(1.1, 1.9)
        r = max(s + rho, s / xmax) + rng.uniform(0.01, 30)
        x = s / r
        poly_I = 0.5 * x**3 / r * np.polyval(H[::-1], x * x)
        poly_dI = -(x**3) / r**2 * np.polyval(K[::-1], x * x)
        assert poly_I == pytest.approx(O.pair_integral(np.array(r), rho, s), rel=3e-7)
        assert poly_dI == pytest.approx(O.pair_integral_dr(np.array(r), rho, s), rel=3e-7)


def test_pair_integral_derivative_fd():
    rng = np.random.default_rng(2)
    for _ in range(300):
        r, rho, s = rng.uniform(0.3, 20), rng.uniform(1.1, 1.9), rng.uniform(0.8, 1.8)
        if min(abs(abs(r - s) - rho), abs(r + s - rho)) < 1e-3:
            continue  # kinks of the max / step
        h = 1e-6
        fd = (O.pair_integral(np.array(r + h), rho, s) - O.pair_integral(np.array(r - h), rho, s)) / (2 * h)
        assert O.pair_integral_dr(np.array(r), rho, s) == pytest.approx(fd, rel=1e-6, abs=1e-9)


# ---------------------------------------------------------------- forces

def _fd_forces(sysd, pos, h=1e-5, **kw):
    coeff = O.cmap_coefficients(sysd["cmap_energy_grids"]) if sysd.get("cmap_torsions") is not None else None
    f = np.zeros_like(pos)
    for i in range(pos.shape[0]):
        for k in range(3):
            xp = pos.copy(); xp[i, k] += h
            xm = pos.copy(); xm[i, k] -= h
            f[i, k] = -(O.energy_terms(sysd, xp, cmap_coeffs=coeff, **kw)["total"]
                        - O.energy_terms(sysd, xm, cmap_coeffs=coeff, **kw)["total"]) / (2 * h)
    return f


@pytest.mark.parametrize("cutoff", [0.0, 6.0])
def test_forces_match_finite_differences(cutoff):
    s = protein(30, 7)
    f = O.forces(s, s["positions"], cutoff=cutoff)
    fd = _fd_forces(s, s["positions"], cutoff=cutoff)
    assert np.sqrt(np.mean((f - fd) ** 2) / np.mean(fd**2)) < 1e-7


def test_harmonic_improper_and_exception_forces_fd():
    s = protein(24, 8)
    rng = np.random.default_rng(3)
    n_imp = len(s["impropers"])
    s["improper_params"] = np.stack([rng.uniform(1, 5, n_imp), rng.uniform(-0.5, 0.5, n_imp)], 1)[:, None, :]
    s["exception_pairs"] = np.array([[0, 10], [2, 20]])
    s["exception_sigmas"] = np.array([2.5, 3.0])
    s["exception_epsilons"] = np.array([0.1, 0.2])
    s["exception_chargeprods"] = np.array([0.05, -0.1])
    s["urey_bradley_bonds"] = s["angles"][:5, [0, 2]]
    s["urey_bradley_params"] = np.stack([np.full(5, 2.5), np.full(5, 20.0)], 1)
    f = O.forces(s, s["positions"])
    fd = _fd_forces(s, s["positions"])
    assert np.sqrt(np.mean((f - fd) ** 2) / np.mean(fd**2)) < 1e-7


def test_newton_and_rigid_motion_invariance():
    s = protein(60, 9)
    pos = s["positions"]
    f = O.forces(s, pos)
    assert np.abs(f.sum(0)).max() < 1e-9 * np.abs(f).max()
    e0 = O.total_energy(s, pos)
    th = 0.7
    R = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    e1 = O.total_energy(s, pos @ R.T + np.array([3.0, -2.0, 5.0]))
    assert e1 == pytest.approx(e0, rel=1e-10)


# ---------------------------------------------------------------- CMAP

def test_bicubic_patch_reproduces_corner_data():
    W = O._bicubic_inverse()
    rng = np.random.default_rng(4)
    rhs = rng.normal(size=16)
    c = (W @ rhs).reshape(4, 4)
    P = lambda a, b: sum(c[p, q] * a**p * b**q for p in range(4) for q in range(4))
    Pa = lambda a, b: sum(p * c[p, q] * a ** max(p - 1, 0) * b**q for p in range(1, 4) for q in range(4))
    for ci, (a, b) in enumerate([(0, 0), (1, 0), (1, 1), (0, 1)]):
        assert P(a, b) == pytest.approx(rhs[ci], abs=1e-12)
        assert Pa(a, b) == pytest.approx(rhs[4 + ci], abs=1e-12)


def test_cmap_interpolant_hits_nodes_and_tracks_smooth_map():
    G = 24
    x = 2 * np.pi * np.arange(G) / G
    phi, psi = np.meshgrid(x, x, indexing="ij")
    Ef = np.sin(phi) * np.cos(2 * psi) + 0.3 * np.cos(phi + psi)
    grid = Ef.T  # stored [psi, phi]
    coeff = O.cmap_coefficients(grid[None])
    # nodes
    e = O.cmap_energy(phi.ravel(), psi.ravel(), np.zeros(G * G, int), coeff)
    assert np.allclose(e, Ef.ravel(), atol=1e-12)
    # between nodes, a smooth periodic map is reproduced closely, values and slopes
    rng = np.random.default_rng(5)
    a, b = rng.uniform(-np.pi, np.pi, 500), rng.uniform(-np.pi, np.pi, 500)
    e, dphi, dpsi = O.cmap_energy(a, b, np.zeros(500, int), coeff, return_grad=True)
    exact = np.sin(a) * np.cos(2 * b) + 0.3 * np.cos(a + b)
    assert np.max(np.abs(e - exact)) < 2e-3
    assert np.max(np.abs(dphi - (np.cos(a) * np.cos(2 * b) - 0.3 * np.sin(a + b)))) < 2e-2


def test_spline_derivatives_of_sinusoid():
    G = 24
    x = np.linspace(0, 2 * np.pi, G + 1)
    y = np.sin(x)
    d = O.spline_derivative_at_nodes(x, y, O.periodic_spline_second_derivs(x, y))
    assert np.max(np.abs(d - np.cos(x[:-1]))) < 1e-4


# ---------------------------------------------------------------- Philox

def test_philox_known_answers():
    """Known-answer vectors published with Random123 for philox4x32-10."""
    assert philox.philox4x32_10((0, 0, 0, 0), (0, 0)) == (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)
    assert philox.philox4x32_10((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF)) == (
        0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)
    assert philox.philox4x32_10((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0)) == (
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)


def test_box_muller_stream_statistics():
    xs = np.array([philox.normal3(12345, s, i) for s in range(40) for i in range(100)]).ravel()
    assert abs(xs.mean()) < 0.05
    assert abs(xs.std() - 1.0) < 0.05


# ------------------------------------------------------- neighbour-list GB form

def test_gb_cutoff_with_complete_list_equals_dense():
    """tests/test_neighbor_list_energy.py:73-100: a list holding every pair gives
    the dense energy. Cutoff above the largest distance -> only the self-term
    form (q^2/B vs q^2/f_ii with the safe_norm eps) differs, by ~1e-13."""
    s = protein(60, 9)
    x = s["positions"]
    rmax = np.max(np.linalg.norm(x[:, None] - x[None], axis=-1))
    dense = O.energy_terms(s, x, cutoff=rmax + 1.0)
    nl = O.energy_terms(s, x, cutoff=rmax + 1.0, gb_cutoff=True)
    assert nl["gb"] == pytest.approx(dense["gb"], rel=1e-11)
    assert np.allclose(nl["_born_radii"], dense["_born_radii"], rtol=1e-13)
    fd = O.forces(s, x, cutoff=rmax + 1.0)
    fn = O.forces(s, x, cutoff=rmax + 1.0, gb_cutoff=True)
    assert np.max(np.abs(fd - fn)) < 1e-9 * np.max(np.abs(fd))


def test_gb_cutoff_separated_ions_are_isolated():
    """Ions farther apart than the cutoff: each has B = R - 0.09 and
    E_GB = pref * sum q^2 / B (generalized_born.py:333-337, 410-414)."""
    s = {"charges": np.array([1.0, -0.5]), "sigmas": np.ones(2), "epsilons": np.zeros(2),
         "radii": np.array([1.7, 2.1])}
    x = np.array([[0.0, 0, 0], [6.0, 0, 0]])
    t = O.energy_terms(s, x, cutoff=5.0, gb_cutoff=True)
    B = np.array([1.61, 2.01])
    assert np.allclose(t["_born_radii"], B, rtol=1e-14)
    assert t["gb"] == pytest.approx(O.gb_prefactor() * np.sum(s["charges"] ** 2 / B), rel=1e-13)
    # the dense form still sees the partner
    assert O.energy_terms(s, x, cutoff=5.0)["gb"] != pytest.approx(t["gb"], rel=1e-6)
    f = O.forces(s, x, cutoff=5.0, gb_cutoff=True)
    assert np.max(np.abs(f)) == 0.0


# ----------------------------------------------------------------------- FIRE

def test_fire_converges_in_a_harmonic_bowl():
    """FIRE on E = 1/2 sum k_i |x_i - c_i|^2 reaches the minimum; dt stays <= dt_max."""
    rng = np.random.default_rng(0)
    c = rng.normal(size=(7, 3))
    k = rng.uniform(5.0, 50.0, size=(7, 1))
    force = lambda x: -k * (x - c)
    x0 = c + rng.normal(size=c.shape)
    x, p, f, dt, alpha, n_pos = O.fire_minimize(force, x0, np.full(7, 12.0), 2000, 0.01, 0.1)
    assert np.max(np.abs(x - c)) < 1e-6
    assert dt <= 0.1 + 1e-15


def test_fire_caps_bound_every_step():
    """With a force cap, no atom moves more than disp_cap in one step, even from a
    steep start (simulate.py:615-640)."""
    rng = np.random.default_rng(1)
    c = rng.normal(size=(5, 3))
    force = lambda x: -1e5 * (x - c)
    x = c + rng.normal(size=c.shape)
    for _ in range(20):
        x_new = O.fire_minimize(force, x, np.ones(5), 1, 0.05, 0.5, force_cap=10.0, disp_cap=0.1)[0]
        assert np.max(np.linalg.norm(x_new - x, axis=1)) <= 0.1 + 1e-12
        x = x_new


def test_capped_descent_bowl_and_cap():
    """Harmonic bowl F = -kappa (x - c): every step moves an atom by
    min(kappa h |x - c|, max_dx) straight toward c, never past it."""
    rng = np.random.default_rng(5)
    c = rng.normal(size=(6, 3))
    x0 = c + rng.normal(size=(6, 3)) * 2.0
    kappa, h, cap = 3.0, 0.002, 0.01
    ff = lambda x: -kappa * (x - c)
    x = x0.copy()
    for _ in range(40):
        x1 = O.capped_descent(ff, x, 1, step_size=h, max_dx=cap)
        d0, d1 = np.linalg.norm(x - c, axis=1), np.linalg.norm(x1 - c, axis=1)
        step = np.minimum(kappa * h * d0, cap)
        assert np.allclose(d0 - d1, step, rtol=1e-9, atol=1e-12)
        x = x1
    assert np.allclose(O.capped_descent(ff, x0, 40, step_size=h, max_dx=cap), x, atol=1e-12)


def test_capped_descent_restraint_balances_a_push():
    """A constant push F0 against the restraint -k (r - r0) settles at
    r0 + F0 / k on the restrained atoms; unrestrained atoms keep moving."""
    x0 = np.zeros((4, 3))
    f0 = np.array([[3.0, 0.0, 0.0], [0.0, -2.0, 0.0], [0.0, 0.0, 1.0], [1.0, 1.0, 1.0]])
    sel = np.array([1.0, 1.0, 1.0, 0.0])
    x = O.capped_descent(lambda x: f0, x0, 3000, step_size=0.002, max_dx=0.01, k_restraint=50.0, restrained=sel)
    assert np.allclose(x[:3], f0[:3] / 50.0, atol=1e-9)
    assert np.allclose(x[3], 3000 * 0.002 * f0[3], rtol=1e-12)


def test_capped_descent_ignores_non_finite_forces():
    x0 = np.ones((3, 3))
    f = np.array([[np.inf, 0, 0], [np.nan, 1.0, 0], [0.0, 0.0, -1.0]])
    x = O.capped_descent(lambda x: f, x0, 1, step_size=0.002, max_dx=0.01)
    assert np.isfinite(x).all()
    assert np.allclose(x - x0, [[0, 0, 0], [0, 0.002, 0], [0, 0, -0.002]])


# ------------------------------------------- independent pins of the formulas
#
# Each check below derives the expected value from the underlying definition
# or the published constants, not from the oracle's own expression, and is
# chosen so that a plausible slip in the oracle (a dropped term, a wrong sign,
# a wrong denominator or exponent) fails it.

def _descreening_quadrature(r, rho, s):
    """HCT descreening integral of sphere j (radius s, centre at distance r > s)
    seen from atom i, outside i's offset sphere rho:
    (1/4pi) * integral over sphere_j minus ball(rho) of |x|^-4 dV, done as a
    1-D quadrature over spherical shells t around atom i. The shell of radius
    t meets sphere j in a cap of solid-angle fraction (1 - cos theta) / 2 with
    cos theta = (t^2 + r^2 - s^2) / (2 t r)."""
    from scipy.integrate import quad

    lo, hi = max(rho, r - s), r + s
    if hi <= lo:
        return 0.0
    f = lambda t: 0.5 * (1.0 - (t * t + r * r - s * s) / (2.0 * t * r)) / (t * t)
    val, err = quad(f, lo, hi, epsabs=0.0, epsrel=1e-13, limit=200)
    return val


def test_pair_integral_equals_shell_quadrature():
    """Both branches of L (partial overlap with the offset sphere, and a sphere
    j entirely outside it) agree with the volume integral to 1e-11."""
    cases = [(2.5, 1.41, 1.275), (2.0, 1.61, 1.1), (3.0, 1.21, 1.5), (6.0, 1.41, 1.3), (1.9, 1.0, 0.95)]
    for r, rho, s in cases:
        assert r > s
        q = _descreening_quadrature(r, rho, s)
        assert O.pair_integral(np.array(r), rho, s) == pytest.approx(q, rel=1e-11), (r, rho, s)


def test_obc2_radii_at_positive_psi_from_quadrature():
    """OBC II rescale (Onufriev, Bashford & Case, Proteins 55:383, 2004: alpha = 1,
    beta = 0.8, gamma = 4.85) on descreening sums taken by quadrature. A buried
    atom inside an icosahedral shell reaches psi = 0.46 (shell atoms 0.2-0.3,
    outer atoms 0.04-0.08), where tanh is far from its argument, so the three
    coefficients, the sign of beta and the 1/R (not 1/rho) denominator all
    decide B_i. Every pair has r > s_j, where prolix's closed form is the
    volume integral."""
    g = (1 + 5**0.5) / 2
    ico = np.array([[0, 1, g], [0, -1, g], [0, 1, -g], [0, -1, -g], [1, g, 0], [-1, g, 0], [1, -g, 0],
                    [-1, -g, 0], [g, 0, 1], [-g, 0, 1], [g, 0, -1], [-g, 0, -1]], float)
    ico /= np.linalg.norm(ico, axis=1, keepdims=True)
    pos = np.vstack([[0.0, 0, 0], ico * 2.3, ico[[0, 3, 5, 8]] * 4.3 + 0.3])
    n = len(pos)
    rng = np.random.default_rng(3)
    R = rng.uniform(1.3, 1.9, n)
    s = R * rng.uniform(0.72, 0.85, n)
    rho = R - 0.09
    I = np.zeros(n)
    for i in range(n):
        for j in range(n):
            if i != j:
                r = float(np.linalg.norm(pos[i] - pos[j]))
                assert r > s[j]
                I[i] += _descreening_quadrature(r, rho[i], s[j])
    psi = rho * I
    assert psi.max() > 0.4 and psi.min() > 0.03
    alpha, beta, gamma = 1.0, 0.8, 4.85
    expect = 1.0 / (1.0 / rho - np.tanh(alpha * psi - beta * psi**2 + gamma * psi**3) / R)
    got = O.born_radii(pos, R, s)
    assert np.allclose(got, expect, rtol=1e-10, atol=0)
    # a flipped beta or a 1/rho denominator moves the buried atom's B by > 1%
    wrong_beta = 1.0 / (1.0 / rho - np.tanh(psi + beta * psi**2 + gamma * psi**3) / R)
    wrong_den = 1.0 / (1.0 / rho - np.tanh(psi - beta * psi**2 + gamma * psi**3) / rho)
    assert abs(wrong_beta[0] / got[0] - 1) > 1e-2 and abs(wrong_den[0] / got[0] - 1) > 1e-2


def test_gb_pair_function_limits():
    """f_GB = sqrt(r^2 + B_i B_j exp(-r^2 / (4 B_i B_j))) (Still et al. 1990) at
    three points fixed by hand, with the Born radii given: widely separated ions
    (f -> r, each unordered pair counted in both orders), r = 2 sqrt(B_i B_j)
    (the exponent is exactly -1) and coincident atoms (f = sqrt(B_i B_j))."""
    q = np.array([0.7, -0.45])
    B = np.array([1.4, 1.9])
    pref = -0.5 * 332.0637 * (1.0 - 1.0 / 78.5)
    self_term = q[0] ** 2 / B[0] + q[1] ** 2 / B[1]
    bb = B[0] * B[1]
    for r, f01 in ((400.0, 400.0), (2.0 * math.sqrt(bb), math.sqrt(bb * (4.0 + math.exp(-1.0)))),
                   (0.0, math.sqrt(bb))):
        pos = np.array([[0.0, 0, 0], [r, 0, 0]])
        e = O.gb_energy(pos, q, B, eps_in=1.0, eps_out=78.5)
        assert e == pytest.approx(pref * (self_term + 2.0 * q[0] * q[1] / f01), rel=1e-11), r


def test_ace_single_ion_by_hand():
    """ACE for one ion of radius 1.5 A (OpenMM's ACE form): B = R - 0.09 = 1.41 A
    for an isolated atom; offset radius in nm 0.141 + 0.009 = 0.15; probe 0.14 nm;
    coefficient 4 pi x 2.25936 kJ/mol/nm^2 (surface tension 0.0054 kcal/mol/A^2);
    kJ -> kcal with prolix's 0.239006."""
    s = {"charges": np.array([0.0]), "sigmas": np.ones(1), "epsilons": np.zeros(1), "radii": np.array([1.5])}
    t = O.energy_terms(s, np.zeros((1, 3)))
    coeff = 4.0 * math.pi * (0.0054 * 4.184 * 100.0)
    expect = coeff * (0.15 + 0.14) ** 2 * (0.15 / 0.141) ** 6 * 0.239006
    assert t["ace"] == pytest.approx(expect, rel=2e-8)
    assert t["gb"] == 0.0  # uncharged


def _dihedral_geometry(phi):
    """Atoms i, j, k, l with j at the origin, k on +z and i on +x; l is i's
    direction turned by phi about the j -> k axis, so the IUPAC dihedral is phi."""
    return np.array([[1.0, 0, 0], [0.0, 0, 0], [0.0, 0, 1.0], [math.cos(phi), math.sin(phi), 1.0]])


def test_dihedral_angle_sign_and_value():
    idx = np.array([[0, 1, 2, 3]])
    for phi in (math.pi / 3, -2.0, 2.9, -0.4):
        assert O.dihedral_angles(_dihedral_geometry(phi), idx)[0] == pytest.approx(phi, abs=1e-7)


def test_periodic_torsion_by_hand():
    """k (1 + cos(n phi - delta)) at phi = pi / 3 (bonded.py:110-141): n = 1,
    delta = pi/3 -> 2k; n = 3, delta = 0 -> 0; n = 2, delta = pi -> k (1 + cos(-pi/3))
    = 1.5 k. Terms of one torsion add."""
    pos = _dihedral_geometry(math.pi / 3)
    idx = np.array([[0, 1, 2, 3]])
    prm = np.array([[[1.0, math.pi / 3, 2.0], [3.0, 0.0, 5.0], [2.0, math.pi, 0.8]]])
    assert O.torsion_energy(pos, idx, prm) == pytest.approx(4.0 + 0.0 + 1.2, abs=1e-6)


def test_harmonic_improper_by_hand():
    """k wrap(phi - phi0)^2 with no 1/2 (bonded.py:168-183), wrapped to [-pi, pi)."""
    idx = np.array([[0, 1, 2, 3]])
    cases = [(math.pi / 3, -math.pi / 2, (5 * math.pi / 6) ** 2),
             (math.pi / 3, 5 * math.pi / 6, (math.pi / 2) ** 2),
             (math.radians(170), math.radians(-170), math.radians(20) ** 2)]
    for phi, phi0, d2 in cases:
        e = O.harmonic_improper_energy(_dihedral_geometry(phi), idx, np.array([[[7.0, phi0]]]))
        assert e == pytest.approx(7.0 * d2, rel=1e-6), (phi, phi0)


def test_exception_pair_by_hand():
    """Pre-scaled 1-4 pair (bonded.py:188-228): LJ is 0 at r = sigma and -eps at
    the minimum r = 2^(1/6) sigma; Coulomb k qq / r."""
    sig, eps, qq = 3.1, 0.15, -0.12
    pairs = np.array([[0, 1]])
    for r, lj in ((sig, 0.0), (2.0 ** (1.0 / 6.0) * sig, -eps)):
        pos = np.array([[0.0, 0, 0], [0.0, r, 0]])
        e = O.exception_energy(pos, pairs, np.array([sig]), np.array([eps]), np.array([qq]))
        assert e == pytest.approx(lj + 332.0637 * qq / r, rel=1e-9, abs=1e-12)


def test_sd_backtrack_in_a_bowl():
    """E = k/2 |x|^2 with every |F| < max_disp / dt_max, so dt = dt_max and the
    candidates are x (1 - f dt k). With dt k = 4 their energies scale as
    (1 - 4 f)^2 = 9, 1, 0, 0.36 for f = 1, 0.5, 0.25, 0.1: the search must
    take f = 0.25, which lands on the minimum exactly."""
    k = 1000.0
    efn = lambda x: 0.5 * k * float(np.sum(x * x))
    ffn = lambda x: -k * x
    x0 = np.array([[0.01, -0.005, 0.002], [0.0, 0.003, -0.004]])
    x1, frac, es, e0 = O.sd_backtrack_step(efn, ffn, x0)
    assert frac == 0.25
    assert np.allclose(x1, 0.0, atol=1e-15)
    assert es == pytest.approx([9 * e0, e0, 0.0, 0.36 * e0], rel=1e-12, abs=1e-18)
    # dt k = 1.5: (1 - 1.5 f)^2 = 0.25, 0.0625, 0.390625, 0.7225 -> f = 0.5
    x1, frac, _, _ = O.sd_backtrack_step(efn, ffn, x0, dt_max=1.5 / k)
    assert frac == 0.5 and np.allclose(x1, 0.25 * x0)
    # a force-limited step: max|F| = k |x|max, dt = 0.1 / max|F|
    xb = np.array([[1.0, 0.0, 0.0], [0.0, 0.5, 0.0]])
    x1, frac, _, _ = O.sd_backtrack_step(efn, ffn, xb)
    assert frac == 1.0 and np.allclose(x1, xb * (1.0 - 0.1 / 1.0))
    # at the minimum nothing is lower than E(r): r stays
    x1, frac, _, _ = O.sd_backtrack_step(efn, ffn, np.zeros((2, 3)))
    assert frac is None and np.array_equal(x1, np.zeros((2, 3)))


def test_fire_stage_best_in_a_bowl():
    """The restrained FIRE sub-stage (batched_simulate.py:813-874) on a harmonic
    bowl: the returned snapshot's RMS force is the lowest one checked (the
    force before each step); with steps that the 0.1 A cap binds, 3 steps move
    no atom more than 0.2 A (the best snapshot is a state before a step); a
    long stage ends near the minimum."""
    k = 50.0
    x0 = np.array([[3.0, -2.0, 1.0], [0.5, 0.4, -2.5]])
    seen = []

    def ff(x):
        f = -k * x
        seen.append(math.sqrt(float(np.sum(f * f)) / len(x)))
        return f

    xb, rms = O.fire_stage_best(ff, x0, np.full(2, 12.0), 300, dt_start=0.01, dt_max=0.1)
    assert rms == min(seen[:-1])
    assert np.max(np.abs(xb)) < 1e-2
    # dt = 0.5: the uncapped first step would move atom 0 by ~3 A
    xb3, _ = O.fire_stage_best(lambda x: -k * x, x0, np.full(2, 12.0), 3, dt_start=0.5, dt_max=0.5)
    moved = np.linalg.norm(xb3 - x0, axis=1)
    assert 0.0 < moved.max() <= 0.2 + 1e-12


def test_lbfgs_reaches_known_minima():
    """L-BFGS restatement: an anisotropic quadratic (minimum 0 at the origin) and
    the Rosenbrock function (minimum 0 at (1, 1)); energies never rise."""
    A = np.diag([1.0, 10.0, 100.0])
    e = lambda x: 0.5 * float(x.reshape(-1) @ A @ x.reshape(-1))
    g = lambda x: (A @ x.reshape(-1)).reshape(x.shape)
    x, tr = O.lbfgs_minimize(e, g, np.array([[1.0, 1.0, 1.0]]), 200, max_stepsize=1.0)
    assert np.max(np.abs(x)) < 1e-6 and np.all(np.diff(tr) <= 0)
    ros = lambda x: (1 - x[0, 0]) ** 2 + 100 * (x[0, 1] - x[0, 0] ** 2) ** 2
    rg = lambda x: np.array([[-2 * (1 - x[0, 0]) - 400 * x[0, 0] * (x[0, 1] - x[0, 0] ** 2),
                              200 * (x[0, 1] - x[0, 0] ** 2)]])
    x, tr = O.lbfgs_minimize(ros, rg, np.array([[-1.2, 1.0]]), 2000, max_stepsize=1.0, max_trials=40)
    assert np.allclose(x, [[1.0, 1.0]], atol=1e-4) and np.all(np.diff(tr) <= 0)
