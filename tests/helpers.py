"""Shared fixtures: seeded systems rounded to float32, with exclusion rows."""

from __future__ import annotations

import numpy as np

from prolix_amd import synth
from prolix_amd.physics.neighbor_list import ExclusionSpec, map_exclusions_to_dense_padded

FLOAT_KEYS = ("positions", "charges", "sigmas", "epsilons", "radii", "scaled_radii", "masses", "bond_params",
              "angle_params", "dihedral_params", "improper_params", "cmap_energy_grids")


def f32_round(sysd: dict) -> dict:
    """Round every float input to float32 (then hold it as float64) so that the
    float64 oracle and the float32 device path see identical numbers."""
    out = dict(sysd)
    for k in FLOAT_KEYS:
        if out.get(k) is not None:
            out[k] = np.asarray(out[k], np.float32).astype(np.float64)
    return out


def with_exclusions(sysd: dict, n_atoms: int) -> dict:
    spec = ExclusionSpec.from_bonds(sysd["bonds"], n_atoms)
    out = dict(sysd)
    out["excl_indices"], out["excl_scales_vdw"], out["excl_scales_elec"] = map_exclusions_to_dense_padded(spec)
    return out


def protein(n_atoms: int, seed: int, radius=None, exclusions=True) -> dict:
    p = synth.make_protein(n_atoms, seed, radius=radius)
    d = p.as_dict()
    if exclusions:
        d = with_exclusions(d, n_atoms)
    d["elements"] = p.elements
    return f32_round(d)


def rel_rms(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.mean((a - b) ** 2) / np.mean(b**2)))


def _round_f32(q):
    """Nearest float32 to the exact rational q (ties to even)."""
    from fractions import Fraction

    f = np.float32(float(q))
    best = None
    for c in (np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))):
        d = abs(Fraction(float(c)) - q)
        key = (d, int(np.asarray(c).view(np.uint32)) & 1)
        if best is None or key < best[0]:
            best = (key, c)
    return best[1]


def kernel_r2_f32(d):
    """r^2 of a float32 difference vector as the HIP pair kernels form it:
    fmaf(dx, dx, fmaf(dy, dy, dz * dz)), each operation rounded once."""
    from fractions import Fraction

    dx, dy, dz = (np.float32(v) for v in d)
    t = np.float32(dz * dz)
    t = _round_f32(Fraction(float(dy)) ** 2 + Fraction(float(t)))
    return _round_f32(Fraction(float(dx)) ** 2 + Fraction(float(t)))


def boundary_pair(rc: float, seed: int = 0):
    """A float32 difference vector whose exact length is >= rc while the
    kernel's float32 r^2 is below (float)(rc^2): the pair is inside the cutoff
    for the device and outside for an exact float64 test."""
    from fractions import Fraction

    rng = np.random.default_rng(seed)
    c2 = np.float32(rc * rc)
    for _ in range(100000):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        d = (u * (rc + 5e-7)).astype(np.float32)
        exact = sum(Fraction(float(v)) ** 2 for v in d)
        if exact >= Fraction(rc) ** 2 and kernel_r2_f32(d) < c2:
            return d
    raise AssertionError("no boundary pair found")
