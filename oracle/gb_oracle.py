"""Float64 NumPy oracle for prolix's GB-OBC2 implicit-solvent energy and forces.

TEST INFRASTRUCTURE ONLY. This module is the checker that the HIP path is
compared against. Only tests, the smoke check and the CPU-baseline leg of the
benchmark may import it; the product path never does.

Each function restates the reference's algorithm from its formulas (file:line
under prolix's ``src/prolix/``):

- ``pair_integral``            physics/generalized_born.py:211-263
- ``born_radii``               physics/generalized_born.py:32-94
- ``gb_energy``                physics/generalized_born.py:97-208 (all ordered
                               pairs incl. i == j; called with mask=1-eye and
                               energy_mask=None from physics/system.py:196-209)
- ``ace_energy``               physics/generalized_born.py:600-648, summed at
                               physics/system.py:234
- ``coulomb_energy``           physics/optimization.py:200-245 (alpha = 0)
- ``lj_energy``                physics/optimization.py:21-68
- ``bond_energy``              physics/bonded.py:20-44 (jax_md simple_spring_bond)
- ``angle_energy``             physics/bonded.py:47-81
- ``dihedral_angles``          physics/bonded.py:84-107
- ``torsion_energy``           physics/bonded.py:110-141
- ``harmonic_improper_energy`` physics/bonded.py:144-185
- ``exception_energy``         physics/bonded.py:188-228
- ``cmap_coefficients`` / ``cmap_energy``   physics/cmap.py:62-389
- ``energy_terms``             physics/system.py:229-248 (term assembly)
- ``gb_cutoff`` options        physics/generalized_born.py:266-415 (neighbour-list
                               Born radii and GB energy, the list taken as every
                               pair closer than the cutoff)
- ``fire_minimize``            simulate.py:615-748 (jax_md fire_descent, capped apply)
- ``capped_descent``           batched_simulate.py:746-787 (per-atom capped SD)
- ``sd_backtrack_step``        physics/simulate.py:141-196 (SD + 4-point backtracking)
- ``fire_stage_best``          batched_simulate.py:813-874 (restrained FIRE sub-stage)
- ``lbfgs_minimize``           batched_simulate.py:886-935 (the L-BFGS polish; jaxopt's own
                               line search is not restated, see the function)

Forces are this module's own analytic derivatives of those energies. Tests pin
them with central finite differences of the energies in float64, Newton's
third law and rigid-motion invariance.

The Born-radius chain rule is F = -(dE/dr|_B + sum_i dE/dB_i * dB_i/dr), the
decomposition prolix documents at physics/generalized_born.py:651-706.

Pinning status: every function here is checked in tests/test_oracle.py by a
value fixed outside this module:

- ``pair_integral``: the HCT shell quadrature of the descreening volume
  integral, and the far-field atanh closed form;
- ``born_radii``: the single ion (B = R - 0.09), and OBC II's published
  alpha / beta / gamma applied to quadrature sums at psi up to 0.46;
- ``gb_energy``: f_GB at r -> infinity, r = 2 sqrt(B_i B_j) and r = 0, with
  the Born radii given;
- ``ace_per_atom``: one ion by hand from the 0.0054 kcal/mol/A^2 surface
  tension;
- ``coulomb_energy``, ``lj_energy``, ``bond_energy``, ``angle_energy``: the
  reference's own closed-form tests;
- ``dihedral_angles``, ``torsion_energy``, ``harmonic_improper_energy``,
  ``exception_energy``: constructed geometries of known dihedral / distance;
- CMAP: node interpolation, C1 periodicity, the analytic spline of a sinusoid;
- ``forces``: central finite differences, Newton's third law, rigid motion;
- the minimisers: a harmonic bowl's known path, cap bounds, a restraint balance;
  the backtracking search: a bowl whose candidate energies are known in
  closed form, so the selection order and the strict comparisons are fixed;
  the FIRE sub-stage: a bowl (steps capped at 0.1 A, the best snapshot's RMS
  force never above any visited); L-BFGS: the exact minimum of an
  anisotropic quadratic and of the Rosenbrock function, with energies that
  never rise.
"""

from __future__ import annotations

import math

import numpy as np

# OBC type II (generalized_born.py:19-21)
ALPHA_OBC = 1.0
BETA_OBC = 0.8
GAMMA_OBC = 4.85
DIELECTRIC_OFFSET = 0.09
ACE_COEFF_KJ_NM = 28.3919551  # generalized_born.py:629
KJ_TO_KCAL = 0.239006  # generalized_born.py:647
COULOMB_CONSTANT = 332.0637  # system.py:135
SAFE_NORM_EPS = 1e-12  # generalized_born.py:27-29
BLOCK = 512


def _row_blocks(n, block=BLOCK):
    for s in range(0, n, block):
        yield s, min(n, s + block)


# ----------------------------------------------------------------------------
# OBC2 pair integral and its radial derivative
# ----------------------------------------------------------------------------

def pair_integral(r, rho_i, s_j):
    """I(r; rho_i, sr_j), generalized_born.py:211-263 (float64, broadcasting)."""
    D = np.abs(r - s_j)
    L = np.maximum(rho_i, D)
    Ls = np.maximum(L, 1e-4)
    U = r + s_j
    Us = np.maximum(U, 1e-4)
    rs = np.maximum(r, 1e-4)
    invL = 1.0 / Ls
    invU = 1.0 / Us
    t1 = 0.5 * (invL - invU)
    t2 = 0.25 * np.log(Ls / Us) / rs
    t3 = 0.125 * (r - s_j * s_j / rs) * (invU * invU - invL * invL)
    return np.where(r + s_j > rho_i, t1 + t2 + t3, 0.0)


def pair_integral_dr(r, rho_i, s_j):
    """dI/dr of ``pair_integral`` with the same branches and clamps."""
    D = np.abs(r - s_j)
    useD = D > rho_i
    L = np.where(useD, D, rho_i)
    Lp = np.where(useD, np.sign(r - s_j), 0.0)
    Ls = np.maximum(L, 1e-4)
    Lp = np.where(L > 1e-4, Lp, 0.0)
    U = r + s_j
    Us = np.maximum(U, 1e-4)
    Up = np.where(U > 1e-4, 1.0, 0.0)
    rs = np.maximum(r, 1e-4)
    rp = np.where(r > 1e-4, 1.0, 0.0)
    invL = 1.0 / Ls
    invU = 1.0 / Us
    lg = np.log(Ls / Us)
    d1 = 0.5 * (-Lp * invL * invL + Up * invU * invU)
    d2 = 0.25 * ((Lp * invL - Up * invU) / rs - lg * rp / (rs * rs))
    d3 = 0.125 * (1.0 + s_j * s_j * rp / (rs * rs)) * (invU * invU - invL * invL)
    d4 = 0.125 * (r - s_j * s_j / rs) * (-2.0 * Up * invU**3 + 2.0 * Lp * invL**3)
    return np.where(r + s_j > rho_i, d1 + d2 + d3 + d4, 0.0)


def _offset_radii(radii, offset, gb_cutoff):
    radii = np.asarray(radii, np.float64)
    if gb_cutoff:  # generalized_born.py:302, 311
        return np.maximum(np.maximum(radii, 1e-6) - offset, 1e-4)
    return radii - offset


def _cut_mask(d, cutoff):
    """Pairs inside the GB cutoff: plain distance r < cutoff (all pairs if cutoff <= 0)."""
    if cutoff <= 0:
        return np.ones(d.shape[:-1], bool)
    return np.sum(d * d, axis=-1) < cutoff * cutoff


def born_sums(pos, radii, scaled_radii, offset=DIELECTRIC_OFFSET, gb_cutoff=0.0):
    """I_i = sum_{j != i} I(r_ij; rho_i, sr_j) with r from safe_norm; with
    ``gb_cutoff`` > 0 only the pairs with r < gb_cutoff contribute."""
    pos = np.asarray(pos, np.float64)
    n = pos.shape[0]
    rho = _offset_radii(radii, offset, gb_cutoff > 0)
    sr = np.asarray(scaled_radii if scaled_radii is not None else radii, np.float64)
    out = np.zeros(n)
    for a, b in _row_blocks(n):
        d = pos[a:b, None, :] - pos[None, :, :]
        r = np.sqrt(np.sum(d * d, axis=-1) + SAFE_NORM_EPS)
        I = np.where(_cut_mask(d, gb_cutoff), pair_integral(r, rho[a:b, None], sr[None, :]), 0.0)
        idx = np.arange(a, b)
        I[idx - a, idx] = 0.0
        out[a:b] = I.sum(axis=1)
    return out


def born_radii(pos, radii, scaled_radii=None, offset=DIELECTRIC_OFFSET, return_aux=False, gb_cutoff=0.0):
    """OBC2 effective Born radii (generalized_born.py:32-94).

    ``gb_cutoff`` > 0 gives the neighbour-list form (generalized_born.py:266-337)
    with the list taken as every pair closer than gb_cutoff: radii clamped as
    there, and 1/B floored at 1e-4.
    """
    radii = np.asarray(radii, np.float64)
    nl = gb_cutoff > 0
    rho = _offset_radii(radii, offset, nl)
    Rd = np.maximum(np.maximum(radii, 1e-6), 1e-4) if nl else radii
    I = born_sums(pos, radii, scaled_radii, offset, gb_cutoff)
    psi = rho * I
    t = ALPHA_OBC * psi - BETA_OBC * psi**2 + GAMMA_OBC * psi**3
    th = np.tanh(t)
    inv_b = 1.0 / rho - th / Rd
    clamped = np.zeros_like(inv_b, bool)
    if nl:
        clamped = inv_b < 1e-4
        inv_b = np.maximum(inv_b, 1e-4)
    B = 1.0 / inv_b
    if not return_aux:
        return B
    dB_dI = B * B * (1.0 - th * th) * (ALPHA_OBC - 2 * BETA_OBC * psi + 3 * GAMMA_OBC * psi**2) * rho / Rd
    dB_dI = np.where(clamped, 0.0, dB_dI)
    return B, {"I": I, "psi": psi, "dB_dI": dB_dI}


# ----------------------------------------------------------------------------
# GB polar energy, ACE non-polar energy
# ----------------------------------------------------------------------------

def gb_prefactor(k_gb=COULOMB_CONSTANT, eps_in=1.0, eps_out=78.5):
    return -0.5 * k_gb * (1.0 / eps_in - 1.0 / eps_out)


def gb_energy(pos, charges, B, k_gb=COULOMB_CONSTANT, eps_in=1.0, eps_out=78.5, gb_cutoff=0.0):
    """E_GB over all ordered pairs including i == j (generalized_born.py:118-208).

    With ``gb_cutoff`` > 0: the neighbour-list form (generalized_born.py:340-415),
    sum over the ordered pairs i != j closer than gb_cutoff plus the self
    term sum_i q_i^2 / B_i.
    """
    pos = np.asarray(pos, np.float64)
    q = np.asarray(charges, np.float64)
    pref = gb_prefactor(k_gb, eps_in, eps_out)
    e = 0.0
    for a, b in _row_blocks(pos.shape[0]):
        d = pos[a:b, None, :] - pos[None, :, :]
        r2 = np.sum(d * d, axis=-1) + SAFE_NORM_EPS  # safe_norm squared
        bb = B[a:b, None] * B[None, :]
        f = np.sqrt(r2 + bb * np.exp(-r2 / (4.0 * bb)))
        t = q[a:b, None] * q[None, :] / f
        if gb_cutoff > 0:
            m = _cut_mask(d, gb_cutoff)
            idx = np.arange(a, b)
            m[idx - a, idx] = False
            t = np.where(m, t, 0.0)
        e += np.sum(t)
    if gb_cutoff > 0:
        e += np.sum(q * q / B)
    return pref * e


def ace_per_atom(radii, B, offset=DIELECTRIC_OFFSET):
    """Per-atom ACE non-polar energy in kcal/mol (generalized_born.py:600-648)."""
    radii = np.asarray(radii, np.float64)
    rn = (radii - offset) / 10.0 + 0.009
    return ACE_COEFF_KJ_NM * (rn + 0.14) ** 2 * (rn / (np.asarray(B) / 10.0)) ** 6 * KJ_TO_KCAL


# ----------------------------------------------------------------------------
# Coulomb and LJ with per-atom exclusion rows
# ----------------------------------------------------------------------------

def dense_scales(excl_idx, excl_scales, n):
    """Pair-scale matrix S[i, j] from (N, C) rows, optimization.py:55-57 / :234-236."""
    S = np.ones((n, n))
    if excl_idx is None or np.size(excl_idx) == 0:
        return S
    excl_idx = np.asarray(excl_idx)
    excl_scales = np.asarray(excl_scales, np.float64)
    acc = np.zeros((n, n))
    hit = np.zeros((n, n), bool)
    for i in range(n):
        for k in range(excl_idx.shape[1]):
            j = int(excl_idx[i, k])
            if 0 <= j < n:
                acc[i, j] += excl_scales[i, k]
                hit[i, j] = True
    return np.where(hit, acc, 1.0)


def _pair_mask(r, cutoff, n_i, a):
    m = r > 0.0
    if cutoff > 0:
        m &= r < cutoff
    idx = np.arange(a, a + n_i)
    m[idx - a, idx] = False
    return m


def coulomb_energy(pos, charges, S_elec, cutoff=0.0, k_c=COULOMB_CONSTANT):
    pos = np.asarray(pos, np.float64)
    q = np.asarray(charges, np.float64)
    e = 0.0
    for a, b in _row_blocks(pos.shape[0]):
        d = pos[a:b, None, :] - pos[None, :, :]
        r = np.sqrt(np.sum(d * d, axis=-1))
        m = _pair_mask(r, cutoff, b - a, a)
        rr = np.where(m, r, 1.0)
        e += np.sum(np.where(m, k_c * q[a:b, None] * q[None, :] * S_elec[a:b] / rr, 0.0))
    return 0.5 * e


def lj_energy(pos, sigmas, epsilons, S_vdw, cutoff=0.0):
    pos = np.asarray(pos, np.float64)
    sig = np.maximum(np.asarray(sigmas, np.float64), 1e-6)
    eps = np.asarray(epsilons, np.float64)
    e = 0.0
    for a, b in _row_blocks(pos.shape[0]):
        d = pos[a:b, None, :] - pos[None, :, :]
        r = np.sqrt(np.sum(d * d, axis=-1))
        m = _pair_mask(r, cutoff, b - a, a)
        rr = np.where(m, r, 1.0)
        s = 0.5 * (sig[a:b, None] + sig[None, :])
        ep = np.sqrt(eps[a:b, None] * eps[None, :])
        x6 = (s / rr) ** 6
        e += np.sum(np.where(m, 4.0 * ep * (x6 * x6 - x6) * S_vdw[a:b], 0.0))
    return 0.5 * e


# ----------------------------------------------------------------------------
# Bonded terms
# ----------------------------------------------------------------------------

def bond_energy(pos, idx, params):
    if len(idx) == 0:
        return 0.0
    d = pos[idx[:, 0]] - pos[idx[:, 1]]
    r = np.sqrt(np.sum(d * d, axis=1))
    return float(np.sum(0.5 * params[:, 1] * (r - params[:, 0]) ** 2))


def _angle_cos(pos, idx):
    v1 = pos[idx[:, 0]] - pos[idx[:, 1]]
    v2 = pos[idx[:, 2]] - pos[idx[:, 1]]
    d1 = np.sqrt(np.sum(v1 * v1, 1))
    d2 = np.sqrt(np.sum(v2 * v2, 1))
    c = np.sum(v1 * v2, 1) / (d1 * d2 + 1e-8)
    return c, v1, v2, d1, d2


def angle_energy(pos, idx, params):
    if len(idx) == 0:
        return 0.0
    c, *_ = _angle_cos(pos, idx)
    th = np.arccos(np.clip(c, -0.999999, 0.999999))
    return float(0.5 * np.sum(params[:, 1] * (th - params[:, 0]) ** 2))


def dihedral_angles(pos, idx):
    """bonded.py:84-107, including the +1e-8 on |b1|."""
    b0 = pos[idx[:, 0]] - pos[idx[:, 1]]
    b1 = pos[idx[:, 2]] - pos[idx[:, 1]]
    b2 = pos[idx[:, 3]] - pos[idx[:, 2]]
    b1n = np.linalg.norm(b1, axis=1, keepdims=True) + 1e-8
    u = b1 / b1n
    v = b0 - np.sum(b0 * u, 1, keepdims=True) * u
    w = b2 - np.sum(b2 * u, 1, keepdims=True) * u
    x = np.sum(v * w, 1)
    y = np.sum(np.cross(u, v) * w, 1)
    return np.arctan2(y, x)


def torsion_energy(pos, idx, params):
    if len(idx) == 0:
        return 0.0
    phi = dihedral_angles(pos, idx)[:, None]
    n, ph, k = params[:, :, 0], params[:, :, 1], params[:, :, 2]
    return float(np.sum(k * (1.0 + np.cos(n * phi - ph))))


def harmonic_improper_energy(pos, idx, params):
    if len(idx) == 0:
        return 0.0
    phi = dihedral_angles(pos, idx)[:, None]
    k, phi0 = params[:, :, 0], params[:, :, 1]
    diff = np.mod(phi - phi0 + np.pi, 2 * np.pi) - np.pi
    return float(np.sum(k * diff**2))


def exception_energy(pos, pairs, sig, eps, qq, k_c=COULOMB_CONSTANT):
    if pairs is None or len(pairs) == 0:
        return 0.0
    d = pos[pairs[:, 0]] - pos[pairs[:, 1]]
    r = np.sqrt(np.sum(d * d, 1) + 1e-12)
    x6 = (sig / r) ** 6
    return float(np.sum(4.0 * eps * (x6 * x6 - x6) + k_c * qq / r))


# ----------------------------------------------------------------------------
# CMAP
# ----------------------------------------------------------------------------

def _bicubic_inverse():
    """16x16 map from corner data to patch coefficients, derived (not tabulated).

    Unknowns c[4p+q] of P(a, b) = sum c[4p+q] a^p b^q. Constraint rows in the
    order [E(4), dE/da(4), dE/db(4), d2E/dadb(4)] at corners (0,0), (1,0),
    (1,1), (0,1) -- the layout of prolix's rhs (cmap.py:264-283).
    """
    corners = [(0, 0), (1, 0), (1, 1), (0, 1)]
    M = np.zeros((16, 16))
    for kind in range(4):
        for ci, (a, b) in enumerate(corners):
            row = kind * 4 + ci
            for p in range(4):
                for q in range(4):
                    if kind == 0:
                        v = a**p * b**q
                    elif kind == 1:
                        v = p * a ** max(p - 1, 0) * b**q if p > 0 else 0.0
                    elif kind == 2:
                        v = q * a**p * b ** max(q - 1, 0) if q > 0 else 0.0
                    else:
                        v = (p * q * a ** max(p - 1, 0) * b ** max(q - 1, 0)) if (p > 0 and q > 0) else 0.0
                    M[row, 4 * p + q] = v
    return np.linalg.inv(M)


def periodic_spline_second_derivs(x, y):
    """cmap.py:62-124: cyclic system, diag 2(h[i-1]+h[i]), off-diagonals as built there."""
    n = len(x) - 1
    h = np.diff(x)
    slopes = np.diff(y) / h
    b = np.zeros(n)
    b[1:] = 6.0 * (slopes[1:] - slopes[:-1])
    b[0] = 6.0 * (slopes[0] - slopes[n - 1])
    diag = 2.0 * (np.roll(h, 1) + h)
    lower = h.copy()
    upper = np.roll(h, -1)
    A = np.diag(diag) + np.diag(lower[1:], -1) + np.diag(upper[:-1], 1)
    A[0, n - 1] = h[n - 1]
    A[n - 1, 0] = h[n - 1]
    m = np.linalg.solve(A, b)
    return np.concatenate([m, m[:1]])


def spline_derivative_at_nodes(x, y, m):
    """cmap.py:127-163 evaluated at t = x[i] (interval i, a = 1, b = 0)."""
    n = len(x) - 1
    i = np.arange(n)
    h = x[i + 1] - x[i]
    return (y[i + 1] - y[i]) / h + h / 6.0 * (-2.0 * m[i] - m[i + 1])


def cmap_coefficients(grids):
    """(M, G, G) raw grids stored [psi, phi] -> (M, G, G, 16) coefficients [phi, psi]."""
    grids = np.asarray(grids, np.float64)
    M_, G, _ = grids.shape
    x = np.linspace(0.0, 2.0 * np.pi, G + 1)
    winv = _bicubic_inverse()
    delta = 2.0 * np.pi / G
    out = np.zeros((M_, G, G, 16))
    for m in range(M_):
        E = grids[m].T  # [phi, psi]

        def deriv_along_axis0(F):
            D = np.zeros_like(F)
            for j in range(G):
                y = np.concatenate([F[:, j], F[:1, j]])
                D[:, j] = spline_derivative_at_nodes(x, y, periodic_spline_second_derivs(x, y))
            return D

        d1 = deriv_along_axis0(E)
        d2 = deriv_along_axis0(E.T).T
        d12 = deriv_along_axis0(d2)
        for i in range(G):
            for j in range(G):
                ni, nj = (i + 1) % G, (j + 1) % G
                cs = [(i, j), (ni, j), (ni, nj), (i, nj)]
                rhs = np.array([E[a, b] for a, b in cs] + [d1[a, b] * delta for a, b in cs]
                               + [d2[a, b] * delta for a, b in cs] + [d12[a, b] * delta * delta for a, b in cs])
                out[m, i, j] = winv @ rhs
    return out


def _cmap_locate(angle, G):
    delta = 2.0 * np.pi / G
    a = np.mod(angle + 2.0 * np.pi, 2.0 * np.pi)
    i = np.floor(a / delta).astype(np.int64)
    frac = a / delta - i
    return i % G, frac


def cmap_energy(phi, psi, map_idx, coeffs, return_grad=False):
    """cmap.py:316-389 with precomputed coefficients."""
    G = coeffs.shape[1]
    i, da = _cmap_locate(phi, G)
    j, db = _cmap_locate(psi, G)
    c = coeffs[map_idx, i, j].reshape(-1, 4, 4)
    pa = np.stack([np.ones_like(da), da, da**2, da**3], 1)
    pb = np.stack([np.ones_like(db), db, db**2, db**3], 1)
    e = np.einsum("np,npq,nq->n", pa, c, pb)
    if not return_grad:
        return e
    dpa = np.stack([np.zeros_like(da), np.ones_like(da), 2 * da, 3 * da**2], 1)
    dpb = np.stack([np.zeros_like(db), np.ones_like(db), 2 * db, 3 * db**2], 1)
    delta = 2.0 * np.pi / G
    dphi = np.einsum("np,npq,nq->n", dpa, c, pb) / delta
    dpsi = np.einsum("np,npq,nq->n", pa, c, dpb) / delta
    return e, dphi, dpsi


# ----------------------------------------------------------------------------
# Assembly
# ----------------------------------------------------------------------------

TERM_NAMES = ("bond", "angle", "urey_bradley", "dihedral", "improper", "cmap", "lj", "exception",
              "gb", "coulomb", "ace")


def _get(sysd, *names):
    for n in names:
        v = sysd.get(n)
        if v is not None and np.size(v) > 0:
            return np.asarray(v)
    return None


def energy_terms(sysd, pos, cutoff=0.0, k_c=COULOMB_CONSTANT, k_gb=COULOMB_CONSTANT, eps_in=1.0,
                 eps_out=78.5, implicit_solvent=True, cmap_coeffs=None, gb_cutoff=False):
    """Per-term energies of make_energy_fn(..., implicit_solvent=True) (system.py:229-248).

    ``cutoff`` applies to LJ and Coulomb; with ``gb_cutoff`` (and cutoff > 0)
    the Born sums and GB pair terms take the neighbour-list form over the
    pairs closer than the cutoff (generalized_born.py:266-415, as called from
    batched_energy.py:999-1010)."""
    pos = np.asarray(pos, np.float64)
    n = pos.shape[0]
    out = dict.fromkeys(TERM_NAMES, 0.0)
    bi = _get(sysd, "bonds")
    if bi is not None:
        out["bond"] = bond_energy(pos, bi.astype(int), np.asarray(sysd["bond_params"], np.float64))
    ub = _get(sysd, "urey_bradley_bonds")
    if ub is not None:
        out["urey_bradley"] = bond_energy(pos, ub.astype(int), np.asarray(sysd["urey_bradley_params"], np.float64))
    ai = _get(sysd, "angles")
    if ai is not None:
        out["angle"] = angle_energy(pos, ai.astype(int), np.asarray(sysd["angle_params"], np.float64))
    di = _get(sysd, "dihedrals", "proper_dihedrals")
    if di is not None:
        out["dihedral"] = torsion_energy(pos, di.astype(int), np.asarray(sysd["dihedral_params"], np.float64))
    ii = _get(sysd, "impropers", "improper_dihedrals")
    if ii is not None:
        ip = np.asarray(sysd["improper_params"], np.float64)
        if ip.shape[-1] == 2:
            out["improper"] = harmonic_improper_energy(pos, ii.astype(int), ip)
        else:
            out["improper"] = torsion_energy(pos, ii.astype(int), ip)
    ct = _get(sysd, "cmap_torsions")
    if ct is not None:
        ct = ct.astype(int)
        coeffs = cmap_coeffs if cmap_coeffs is not None else cmap_coefficients(sysd["cmap_energy_grids"])
        phi = dihedral_angles(pos, ct[:, 0:4])
        psi = dihedral_angles(pos, ct[:, 1:5])
        out["cmap"] = float(np.sum(cmap_energy(phi, psi, np.asarray(sysd["cmap_indices"]).astype(int), coeffs)))
    S_v = dense_scales(sysd.get("excl_indices"), sysd.get("excl_scales_vdw"), n)
    S_e = dense_scales(sysd.get("excl_indices"), sysd.get("excl_scales_elec"), n)
    out["lj"] = lj_energy(pos, sysd["sigmas"], sysd["epsilons"], S_v, cutoff)
    ep = _get(sysd, "exception_pairs")
    if ep is not None:
        out["exception"] = exception_energy(pos, ep.astype(int), np.asarray(sysd["exception_sigmas"], np.float64),
                                            np.asarray(sysd["exception_epsilons"], np.float64),
                                            np.asarray(sysd["exception_chargeprods"], np.float64), k_c)
    out["coulomb"] = coulomb_energy(pos, sysd["charges"], S_e, cutoff, k_c)
    if implicit_solvent:
        radii = np.asarray(sysd["radii"], np.float64)
        gcut = float(cutoff) if gb_cutoff else 0.0
        B = born_radii(pos, radii, sysd.get("scaled_radii"), gb_cutoff=gcut)
        out["gb"] = gb_energy(pos, sysd["charges"], B, k_gb, eps_in, eps_out, gb_cutoff=gcut)
        out["ace"] = float(np.sum(ace_per_atom(radii, B)))
        out["_born_radii"] = B
    out["total"] = float(sum(v for k, v in out.items() if k in TERM_NAMES))
    return out


def total_energy(sysd, pos, **kw):
    return energy_terms(sysd, pos, **kw)["total"]


# ----------------------------------------------------------------------------
# Analytic forces
# ----------------------------------------------------------------------------

def _dihedral_grad(pos, idx):
    """d phi / d x for the four atoms (Blondel-Karplus form), shape (n, 4, 3)."""
    p0, p1, p2, p3 = (pos[idx[:, k]] for k in range(4))
    F = p0 - p1
    G = p1 - p2
    H = p3 - p2
    A = np.cross(F, G)
    Bv = np.cross(H, G)
    A2 = np.sum(A * A, 1)
    B2 = np.sum(Bv * Bv, 1)
    Gn = np.sqrt(np.sum(G * G, 1))
    FG = np.sum(F * G, 1)
    HG = np.sum(H * G, 1)
    g0 = -(Gn / A2)[:, None] * A
    g3 = (Gn / B2)[:, None] * Bv
    g1 = (Gn / A2)[:, None] * A + (FG / (A2 * Gn))[:, None] * A - (HG / (B2 * Gn))[:, None] * Bv
    g2 = -(Gn / B2)[:, None] * Bv - (FG / (A2 * Gn))[:, None] * A + (HG / (B2 * Gn))[:, None] * Bv
    return np.stack([g0, g1, g2, g3], axis=1)


def _scatter_torsion(f, idx, dE_dphi, grad):
    for k in range(4):
        np.add.at(f, idx[:, k], -dE_dphi[:, None] * grad[:, k])


def forces(sysd, pos, cutoff=0.0, k_c=COULOMB_CONSTANT, k_gb=COULOMB_CONSTANT, eps_in=1.0, eps_out=78.5,
           implicit_solvent=True, cmap_coeffs=None, return_aux=False, gb_cutoff=False):
    """F = -dE/dx of ``energy_terms`` (float64)."""
    pos = np.asarray(pos, np.float64)
    n = pos.shape[0]
    f = np.zeros((n, 3))
    # bonds / UB
    for key, pkey in (("bonds", "bond_params"), ("urey_bradley_bonds", "urey_bradley_params")):
        bi = _get(sysd, key)
        if bi is None:
            continue
        bi = bi.astype(int)
        p = np.asarray(sysd[pkey], np.float64)
        d = pos[bi[:, 0]] - pos[bi[:, 1]]
        r = np.sqrt(np.sum(d * d, 1))
        g = (p[:, 1] * (r - p[:, 0]) / r)[:, None] * d  # dE/dx_0
        np.add.at(f, bi[:, 0], -g)
        np.add.at(f, bi[:, 1], g)
    # angles
    ai = _get(sysd, "angles")
    if ai is not None:
        ai = ai.astype(int)
        p = np.asarray(sysd["angle_params"], np.float64)
        c, v1, v2, d1, d2 = _angle_cos(pos, ai)
        cc = np.clip(c, -0.999999, 0.999999)
        th = np.arccos(cc)
        inside = (c > -0.999999) & (c < 0.999999)
        dE_dc = np.where(inside, p[:, 1] * (th - p[:, 0]) * (-1.0 / np.sqrt(1.0 - cc * cc)), 0.0)
        Q = d1 * d2 + 1e-8
        P = np.sum(v1 * v2, 1)
        dc_d0 = v2 / Q[:, None] - (P / Q**2 * d2 / d1)[:, None] * v1
        dc_d2 = v1 / Q[:, None] - (P / Q**2 * d1 / d2)[:, None] * v2
        g0 = dE_dc[:, None] * dc_d0
        g2 = dE_dc[:, None] * dc_d2
        np.add.at(f, ai[:, 0], -g0)
        np.add.at(f, ai[:, 2], -g2)
        np.add.at(f, ai[:, 1], g0 + g2)
    # torsions
    for key, pkey, alt in (("dihedrals", "dihedral_params", "proper_dihedrals"),
                           ("impropers", "improper_params", "improper_dihedrals")):
        ti = _get(sysd, key, alt)
        if ti is None:
            continue
        ti = ti.astype(int)
        p = np.asarray(sysd[pkey], np.float64)
        phi = dihedral_angles(pos, ti)
        if p.shape[-1] == 2:
            k, phi0 = p[:, :, 0], p[:, :, 1]
            diff = np.mod(phi[:, None] - phi0 + np.pi, 2 * np.pi) - np.pi
            dE = np.sum(2.0 * k * diff, 1)
        else:
            nn, ph, k = p[:, :, 0], p[:, :, 1], p[:, :, 2]
            dE = np.sum(-k * nn * np.sin(nn * phi[:, None] - ph), 1)
        _scatter_torsion(f, ti, dE, _dihedral_grad(pos, ti))
    # CMAP
    ct = _get(sysd, "cmap_torsions")
    if ct is not None:
        ct = ct.astype(int)
        coeffs = cmap_coeffs if cmap_coeffs is not None else cmap_coefficients(sysd["cmap_energy_grids"])
        phi = dihedral_angles(pos, ct[:, 0:4])
        psi = dihedral_angles(pos, ct[:, 1:5])
        _, dphi, dpsi = cmap_energy(phi, psi, np.asarray(sysd["cmap_indices"]).astype(int), coeffs, True)
        _scatter_torsion(f, ct[:, 0:4], dphi, _dihedral_grad(pos, ct[:, 0:4]))
        _scatter_torsion(f, ct[:, 1:5], dpsi, _dihedral_grad(pos, ct[:, 1:5]))
    # exceptions
    ep = _get(sysd, "exception_pairs")
    if ep is not None:
        ep = ep.astype(int)
        sg = np.asarray(sysd["exception_sigmas"], np.float64)
        es = np.asarray(sysd["exception_epsilons"], np.float64)
        qq = np.asarray(sysd["exception_chargeprods"], np.float64)
        d = pos[ep[:, 0]] - pos[ep[:, 1]]
        r = np.sqrt(np.sum(d * d, 1) + 1e-12)
        x6 = (sg / r) ** 6
        dEdr = 4.0 * es * (-12.0 * x6 * x6 + 6.0 * x6) / r - k_c * qq / (r * r)
        g = (dEdr / r)[:, None] * d
        np.add.at(f, ep[:, 0], -g)
        np.add.at(f, ep[:, 1], g)
    # LJ + Coulomb (exact gradient of 0.5 * sum_{i != j} S[i, j] * pair)
    S_v = dense_scales(sysd.get("excl_indices"), sysd.get("excl_scales_vdw"), n)
    S_e = dense_scales(sysd.get("excl_indices"), sysd.get("excl_scales_elec"), n)
    S_v = 0.5 * (S_v + S_v.T)
    S_e = 0.5 * (S_e + S_e.T)
    q = np.asarray(sysd["charges"], np.float64)
    sig = np.maximum(np.asarray(sysd["sigmas"], np.float64), 1e-6)
    eps = np.asarray(sysd["epsilons"], np.float64)
    for a, b in _row_blocks(n):
        d = pos[a:b, None, :] - pos[None, :, :]
        r = np.sqrt(np.sum(d * d, axis=-1))
        m = _pair_mask(r, cutoff, b - a, a)
        rr = np.where(m, r, 1.0)
        s = 0.5 * (sig[a:b, None] + sig[None, :])
        e_ = np.sqrt(eps[a:b, None] * eps[None, :])
        x6 = (s / rr) ** 6
        # -dE/dr / r  for the pair (unordered energy counted once after symmetrisation)
        fr = 4.0 * e_ * (12.0 * x6 * x6 - 6.0 * x6) / rr**2 * S_v[a:b]
        fr += k_c * q[a:b, None] * q[None, :] * S_e[a:b] / rr**3
        fr = np.where(m, fr, 0.0)
        f[a:b] += np.sum(fr[:, :, None] * d, axis=1)
    aux = {}
    if implicit_solvent:
        radii = np.asarray(sysd["radii"], np.float64)
        sr = np.asarray(sysd.get("scaled_radii") if sysd.get("scaled_radii") is not None else radii, np.float64)
        gcut = float(cutoff) if gb_cutoff else 0.0
        rho = _offset_radii(radii, DIELECTRIC_OFFSET, gcut > 0)
        B, bx = born_radii(pos, radii, sr, return_aux=True, gb_cutoff=gcut)
        pref = gb_prefactor(k_gb, eps_in, eps_out)
        dEdB = np.zeros(n)
        if gcut > 0:  # self term q^2 / B
            dEdB += -pref * q * q / (B * B)
        else:  # self term, f_ii^2 = eps + B^2 exp(-eps / (4 B^2))
            e1 = np.exp(-SAFE_NORM_EPS / (4.0 * B * B))
            fii = np.sqrt(SAFE_NORM_EPS + B * B * e1)
            dEdB += -pref * q * q / (2.0 * fii**3) * e1 * (2.0 * B + SAFE_NORM_EPS / (2.0 * B))
        # ACE
        ace = ace_per_atom(radii, B)
        dEdB += -6.0 * ace / B
        for a, b in _row_blocks(n):
            d = pos[a:b, None, :] - pos[None, :, :]
            r2 = np.sum(d * d, axis=-1) + SAFE_NORM_EPS
            r = np.sqrt(r2)
            bb = B[a:b, None] * B[None, :]
            ex = np.exp(-r2 / (4.0 * bb))
            fgb = np.sqrt(r2 + bb * ex)
            qq = q[a:b, None] * q[None, :]
            idx = np.arange(a, b)
            off = _cut_mask(d, gcut)
            off[idx - a, idx] = False
            # direct part at fixed B: E = pref * qq / f over both orders -> 2 per unordered pair
            dEdr_over_r = np.where(off, 2.0 * pref * qq * (-(1.0 - ex / 4.0) / fgb**3), 0.0)
            f[a:b] -= np.sum(dEdr_over_r[:, :, None] * d, axis=1)
            # dE/dB_i from ordered pairs (i, j) and (j, i)
            dfb = B[None, :] * ex * (1.0 + r2 / (4.0 * bb))
            dEdB[a:b] += np.sum(np.where(off, 2.0 * pref * qq * (-0.5) / fgb**3 * dfb, 0.0), axis=1)
        c = dEdB * bx["dB_dI"]
        for a, b in _row_blocks(n):
            d = pos[a:b, None, :] - pos[None, :, :]
            r = np.sqrt(np.sum(d * d, axis=-1) + SAFE_NORM_EPS)
            idx = np.arange(a, b)
            dIij = pair_integral_dr(r, rho[a:b, None], sr[None, :])
            dIji = pair_integral_dr(r, rho[None, :], sr[a:b, None])
            g = c[a:b, None] * dIij + c[None, :] * dIji
            g = np.where(_cut_mask(d, gcut), g, 0.0)
            g[idx - a, idx] = 0.0
            f[a:b] -= np.sum((g / r)[:, :, None] * d, axis=1)
        aux = {"born_radii": B, "dE_dB": dEdB, "chain": c, **bx}
    if return_aux:
        return f, aux
    return f


# ----------------------------------------------------------------------------
# Integrator (BAOAB, physics/settle.py:2064-2087 composed as jax_md nvt_langevin)
# ----------------------------------------------------------------------------

AKMA_TIME_UNIT_FS = 48.88821291839  # simulate.py:56
BOLTZMANN_KCAL = 0.0019872041  # simulate.py:57


def baoab_step(pos, mom, force, mass, dt, kT, gamma, noise, force_fn):
    """One B-A-O-A-force-B step in AKMA units; ``noise`` is the (N, 3) normal draw."""
    m = np.asarray(mass, np.float64)[:, None]
    mom = mom + 0.5 * dt * force
    pos = pos + 0.5 * dt * mom / m
    c1 = math.exp(-gamma * dt)
    c2 = math.sqrt(1.0 - c1 * c1)
    mom = c1 * mom + c2 * np.sqrt(m * kT) * noise
    pos = pos + 0.5 * dt * mom / m
    force = force_fn(pos)
    mom = mom + 0.5 * dt * force
    return pos, mom, force


# ----------------------------------------------------------------------------
# FIRE minimisation (jax_md.minimize.fire_descent as staged by prolix's
# simulate.py:615-748), float64, one system
# ----------------------------------------------------------------------------

def fire_minimize(force_fn, pos, mass, steps, dt_start, dt_max, force_cap=None, disp_cap=0.1, n_min=5,
                  f_inc=1.1, f_dec=0.5, alpha_start=0.1, f_alpha=0.99):
    """Returns (positions, momenta, forces, dt, alpha, n_pos) after ``steps`` steps."""
    x = np.asarray(pos, np.float64).copy()
    m = np.asarray(mass, np.float64).reshape(-1, 1)
    p = np.zeros_like(x)
    f = force_fn(x)
    dt, alpha, n_pos = float(dt_start), float(alpha_start), 0
    for _ in range(int(steps)):
        x0 = x
        if force_cap is not None:  # simulate.py:618-623
            fn = np.linalg.norm(f, axis=1, keepdims=True)
            f = f * np.minimum(1.0, force_cap / (fn + 1e-8))
            f = np.where(np.isfinite(f), f, 0.0)
        p = p + 0.5 * dt * f
        x = x0 + dt * p / m
        f = force_fn(x)
        p = p + 0.5 * dt * f
        f_norm = math.sqrt(float(np.sum(f * f)) + 1e-6)
        p_norm = math.sqrt(float(np.sum(p * p)))
        fdotp = float(np.sum(f * p))
        p = p + alpha * (f * p_norm / f_norm - p)
        n_pos = n_pos + 1 if fdotp >= 0 else 0
        if fdotp > 0 and n_pos > n_min:
            dt = min(dt * f_inc, dt_max)
            alpha = alpha * f_alpha
        if fdotp < 0:
            dt = dt * f_dec
            alpha = alpha_start
            p = np.zeros_like(p)
        if force_cap is not None and disp_cap is not None:  # simulate.py:628-638
            dr = x - x0
            sc = np.minimum(1.0, disp_cap / (np.linalg.norm(dr, axis=1, keepdims=True) + 1e-8))
            x = x0 + dr * sc
            p = p * sc
    return x, p, f, dt, alpha, n_pos


# ----------------------------------------------------------------------------
# Capped steepest descent: the first stage of the batched minimiser
# (batched_simulate.py:746-787), float64, one system
# ----------------------------------------------------------------------------

def capped_descent(force_fn, pos, steps, step_size=0.002, max_dx=0.01, k_restraint=0.0, restrained=None):
    """Positions after ``steps`` steps of dr = F * step_size, |dr| <= max_dx per atom
    (|dr| = sqrt(dr.dr + 1e-30)), with the optional restraint -k (r - r0) on the
    flagged atoms; non-finite forces count as 0, and non-finite final positions
    roll back to the start."""
    x0 = np.asarray(pos, np.float64).copy()
    x = x0.copy()
    sel = None
    if k_restraint > 0.0:
        sel = np.ones(len(x)) if restrained is None else np.asarray(restrained, np.float64).reshape(-1)
    for _ in range(int(steps)):
        f = force_fn(x)
        if sel is not None:
            f = f - k_restraint * (x - x0) * sel[:, None]
        f = np.where(np.isfinite(f), f, 0.0)
        dr = f * step_size
        mag = np.sqrt(np.sum(dr * dr, axis=1, keepdims=True) + 1e-30)
        x = x + dr * np.minimum(1.0, max_dx / mag)
    return x if np.isfinite(x).all() else x0


# ----------------------------------------------------------------------------
# Steepest descent with a 4-point backtracking search (physics/simulate.py:141-196),
# float64, one system
# ----------------------------------------------------------------------------

SD_FRACTIONS = (1.0, 0.5, 0.25, 0.1)


def sd_backtrack_step(energy_fn, force_fn, pos, dt_max=4e-3, max_displacement=0.1):
    """One iteration of physics/simulate.py:174-191: dt = min(max_disp / max|F|, dt_max)
    with max|F| + 1e-8, candidates r + {1, 0.5, 0.25, 0.1} dt F; the 0.1
    candidate is the start and 0.25, 0.5, 1 replace it when strictly lower; the
    best replaces r when strictly below E(r). Returns (positions, chosen fraction
    or None, the four candidate energies, E(r))."""
    x = np.asarray(pos, np.float64)
    e0 = float(energy_fn(x))
    f = force_fn(x)
    fmax = float(np.max(np.linalg.norm(f, axis=1))) + 1e-8
    dt = min(max_displacement / fmax, dt_max)
    cands = [x + frac * dt * f for frac in SD_FRACTIONS]
    es = [float(energy_fn(c)) for c in cands]
    best = 3
    for k in (2, 1, 0):
        if es[k] < es[best]:
            best = k
    if es[best] < e0:
        return cands[best], SD_FRACTIONS[best], es, e0
    return x, None, es, e0


# ----------------------------------------------------------------------------
# The batched minimiser's restrained FIRE sub-stage and L-BFGS polish
# (batched_simulate.py:801-935), float64, one system
# ----------------------------------------------------------------------------

def fire_stage_best(force_fn, pos, mass, steps, dt_start=1e-4, dt_max=1e-3, disp_cap=0.1, n_min=5, f_inc=1.1,
                    f_dec=0.5, alpha_start=0.1, f_alpha=0.99):
    """One FIRE sub-stage (batched_simulate.py:813-874): non-finite force
    components -> 0 (no magnitude cap), the per-step displacement capped at
    disp_cap with |dr| = sqrt(dr.dr + 1e-30) and the momentum scaled alike;
    the positions with the lowest RMS force seen before a step are returned
    with that RMS force."""
    x = np.asarray(pos, np.float64).copy()
    m = np.asarray(mass, np.float64).reshape(-1, 1)
    p = np.zeros_like(x)
    guard = lambda f: np.where(np.isfinite(f), f, 0.0)
    f = guard(force_fn(x))
    dt, alpha, n_pos = float(dt_start), float(alpha_start), 0
    best_x, best_rms = x.copy(), 1e10
    for _ in range(int(steps)):
        rms = math.sqrt(float(np.sum(f * f)) / len(x))
        if rms < best_rms:
            best_x, best_rms = x.copy(), rms
        x0 = x
        p = p + 0.5 * dt * f
        x = x0 + dt * p / m
        f = guard(force_fn(x))
        p = p + 0.5 * dt * f
        f_norm = math.sqrt(float(np.sum(f * f)) + 1e-6)
        p_norm = math.sqrt(float(np.sum(p * p)))
        fdotp = float(np.sum(f * p))
        p = p + alpha * (f * p_norm / f_norm - p)
        n_pos = n_pos + 1 if fdotp >= 0 else 0
        if fdotp > 0 and n_pos > n_min:
            dt = min(dt * f_inc, dt_max)
            alpha = alpha * f_alpha
        if fdotp < 0:
            dt = dt * f_dec
            alpha = alpha_start
            p = np.zeros_like(p)
        dr = x - x0
        sc = np.minimum(1.0, disp_cap / np.sqrt(np.sum(dr * dr, axis=1, keepdims=True) + 1e-30))
        x = x0 + dr * sc
        p = p * sc
    return best_x, best_rms


def lbfgs_minimize(energy_fn, grad_fn, pos, steps, history=10, max_stepsize=0.1, tol=1e-6, c1=1e-4, shrink=0.5,
                   max_trials=12):
    """L-BFGS with a backtracking Armijo search, the method the device runs for
    the batched minimiser's last stage (batched_simulate.py:886-935 runs
    jaxopt.LBFGS, history 10, max_stepsize 0.1): two-loop recursion (Nocedal &
    Wright, Alg. 7.4) with H0 = (s.y / y.y) I over the last ``history`` pairs
    with s.y > 0; trial steps t = max_stepsize * shrink^k; a failed search or a
    gradient norm <= tol leaves x unchanged for that iteration; non-finite
    energies count as 1e10 and non-finite gradient components as 0. Returns
    (positions, energies per iteration)."""
    fe = lambda y: (lambda e: e if np.isfinite(e) else 1e10)(float(energy_fn(y)))
    fg = lambda y: (lambda g: np.where(np.isfinite(g), g, 0.0))(np.asarray(grad_fn(y), np.float64))
    x = np.asarray(pos, np.float64).copy()
    e, g = fe(x), fg(x)
    S, Y = [], []
    trace = [e]
    for _ in range(int(steps)):
        if math.sqrt(float(np.sum(g * g))) <= tol:
            break
        q = g.copy()
        alphas = []
        for s_, y_ in reversed(list(zip(S, Y))):
            a = float(np.sum(s_ * q)) / float(np.sum(s_ * y_))
            alphas.append(a)
            q = q - a * y_
        gamma = 1.0
        if S:
            sy, yy = float(np.sum(S[-1] * Y[-1])), float(np.sum(Y[-1] * Y[-1]))
            if sy > 0 and yy > 0:
                gamma = sy / yy
        r = gamma * q
        for (s_, y_), a in zip(zip(S, Y), reversed(alphas)):
            b = float(np.sum(y_ * r)) / float(np.sum(s_ * y_))
            r = r + (a - b) * s_
        d = -r
        gd = float(np.sum(g * d))
        if not gd < 0:
            d, gd = -g, -float(np.sum(g * g))
        t, step, e_new = float(max_stepsize), 0.0, e
        for _k in range(int(max_trials)):
            et = fe(x + t * d)
            if et <= e + c1 * t * gd:
                step, e_new = t, et
                break
            t *= shrink
        if step > 0:
            x_new = x + step * d
            g_new = fg(x_new)
            s_v, y_v = x_new - x, g_new - g
            if float(np.sum(s_v * y_v)) > 1e-300:
                S.append(s_v)
                Y.append(y_v)
                if len(S) > history:
                    S.pop(0)
                    Y.pop(0)
            x, g, e = x_new, g_new, e_new
        trace.append(e)
    return x, np.array(trace)
