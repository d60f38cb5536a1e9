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
