"""Sparse exclusion specification and its per-atom (N, C) row form.

Mirrors prolix's ``physics/neighbor_list.py``:

- ``ExclusionSpec`` / ``ExclusionSpec.from_protein``  (:25-142)
- ``max_exclusion_slots_needed``                     (:208-233)
- ``map_exclusions_to_dense_padded``                 (:236-303)

One deliberate difference: the reference silently truncates an atom's row at
``max_exclusions`` (:297); here that raises ``ValueError`` instead, because a
dropped exclusion changes the physics without any signal.
"""

from __future__ import annotations

import dataclasses

import numpy as np

from prolix_amd.utils import topology

COULOMB14_SCALE = 0.83333333
LJ14_SCALE = 0.5


@dataclasses.dataclass(frozen=True)
class ExclusionSpec:
    """1-2/1-3 pairs (fully excluded), 1-4 pairs (scaled) and explicit exceptions."""

    idx_12_13: np.ndarray  # (M1, 2) int32
    idx_14: np.ndarray  # (M2, 2) int32
    scale_14_elec: float
    scale_14_vdw: float
    n_atoms: int
    exception_pairs: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 2), np.int32))
    exception_sigmas: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0,), np.float32))
    exception_epsilons: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0,), np.float32))
    exception_chargeprods: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0,), np.float32))

    @classmethod
    def from_bonds(cls, bonds, n_atoms: int, coulomb14scale: float = COULOMB14_SCALE,
                   lj14scale: float = LJ14_SCALE) -> "ExclusionSpec":
        """Topology-derived spec (the bonds-only branch of from_protein, :58-142)."""
        ex = topology.find_bonded_exclusions(bonds, n_atoms)
        return cls(
            idx_12_13=np.concatenate([ex.idx_12, ex.idx_13], axis=0).astype(np.int32),
            idx_14=ex.idx_14.astype(np.int32),
            scale_14_elec=float(coulomb14scale),
            scale_14_vdw=float(lj14scale),
            n_atoms=int(n_atoms),
        )

    @classmethod
    def from_protein(cls, protein, coulomb14scale: float | None = None,
                     lj14scale: float | None = None) -> "ExclusionSpec":
        """Build from an object with ``charges`` and ``bonds`` (and optional 1-4 exceptions).

        With ``pairs_14`` and ``exception_14_params`` present, every listed 1-4
        pair is fully excluded from the main kernels and its energy comes from
        the exception term instead (:96-131).
        """
        charges = getattr(protein, "charges", None)
        if charges is None:
            raise ValueError("Protein must have charges to build ExclusionSpec")
        n = len(charges)
        c14 = coulomb14scale if coulomb14scale is not None else (
            getattr(protein, "coulomb14scale", None) or COULOMB14_SCALE)
        l14 = lj14scale if lj14scale is not None else (getattr(protein, "lj14scale", None) or LJ14_SCALE)
        bonds = getattr(protein, "bonds", None)
        if bonds is None or len(bonds) == 0:
            return cls(np.zeros((0, 2), np.int32), np.zeros((0, 2), np.int32), float(c14), float(l14), n)
        spec = cls.from_bonds(bonds, n, c14, l14)
        pairs14 = getattr(protein, "pairs_14", None)
        params14 = getattr(protein, "exception_14_params", None)
        if pairs14 is not None and len(pairs14) > 0 and params14 is not None and len(params14) > 0:
            pairs = np.asarray(pairs14, np.int32)
            prm = np.asarray(params14, np.float32)
            return dataclasses.replace(
                spec,
                idx_12_13=np.concatenate([spec.idx_12_13, pairs], axis=0),
                idx_14=np.zeros((0, 2), np.int32),
                exception_pairs=pairs,
                exception_chargeprods=prm[:, 0].copy(),
                exception_sigmas=prm[:, 1].copy(),
                exception_epsilons=prm[:, 2].copy(),
            )
        return spec


def _atom_entries(spec: ExclusionSpec):
    n = spec.n_atoms
    rows = [[] for _ in range(n)]
    for i, j in np.asarray(spec.idx_12_13).reshape(-1, 2):
        if i < n:
            rows[i].append((int(j), 0.0, 0.0))
        if j < n:
            rows[j].append((int(i), 0.0, 0.0))
    for i, j in np.asarray(spec.idx_14).reshape(-1, 2):
        if i < n:
            rows[i].append((int(j), spec.scale_14_vdw, spec.scale_14_elec))
        if j < n:
            rows[j].append((int(i), spec.scale_14_vdw, spec.scale_14_elec))
    return rows


def max_exclusion_slots_needed(spec: ExclusionSpec) -> int:
    """Longest per-atom exclusion row (neighbor_list.py:208-233)."""
    rows = _atom_entries(spec)
    return max((len(r) for r in rows), default=0)


def map_exclusions_to_dense_padded(spec: ExclusionSpec, max_exclusions: int = 32):
    """(N, C) int32 partner rows (-1 padded, sorted) with vdw / elec scales (default 1)."""
    n = spec.n_atoms
    rows = _atom_entries(spec)
    need = max((len(r) for r in rows), default=0)
    if need > max_exclusions:
        raise ValueError(
            f"map_exclusions_to_dense_padded: an atom needs {need} exclusion slots but "
            f"max_exclusions={max_exclusions}; entries would be dropped")
    idx = np.full((n, max_exclusions), -1, dtype=np.int32)
    sv = np.ones((n, max_exclusions), dtype=np.float32)
    se = np.ones((n, max_exclusions), dtype=np.float32)
    for i, ent in enumerate(rows):
        ent.sort(key=lambda x: x[0])
        for k, (j, v, e) in enumerate(ent):
            idx[i, k] = j
            sv[i, k] = v
            se[i, k] = e
    return idx, sv, se
