"""Sparse exclusion specification and its per-atom (N, C) row form.

Mirrors prolix's ``physics/neighbor_list.py``:

- ``ExclusionSpec`` / ``ExclusionSpec.from_protein``  (:25-142)
- ``max_exclusion_slots_needed``                     (:208-233)
- ``map_exclusions_to_dense_padded``                 (:236-303)
- ``make_neighbor_list_fn``                          (:155-172)
- ``get_neighbor_exclusion_scales``                  (:382-448)
- ``compute_nl_capacity``                            (:451-530)

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


# ---------------------------------------------------------------------------
# Neighbour-list object and its helpers (neighbor_list.py:155-172, 382-530)
# ---------------------------------------------------------------------------

def get_neighbor_exclusion_scales(excl_indices, excl_scales_vdw, excl_scales_elec, neighbor_idx):
    """(scale_vdw, scale_elec), each (N, K): the exclusion scales of every listed
    neighbour, 1 where the pair is not in the atom's exclusion row
    (neighbor_list.py:382-448: columns m = 0..M-1 in order, a later matching
    column overwriting an earlier one). Works on NumPy arrays or torch tensors
    (on their device)."""
    is_torch = hasattr(neighbor_idx, "device") and hasattr(neighbor_idx, "dtype") and not isinstance(
        neighbor_idx, np.ndarray)
    if is_torch:
        import torch

        dev = neighbor_idx.device
        ex = torch.as_tensor(excl_indices, device=dev)
        sv = torch.as_tensor(excl_scales_vdw, dtype=torch.float32, device=dev)
        se = torch.as_tensor(excl_scales_elec, dtype=torch.float32, device=dev)
        out_v = torch.ones(neighbor_idx.shape, dtype=torch.float32, device=dev)
        out_e = torch.ones_like(out_v)
        for m in range(ex.shape[1]):
            col = ex[:, m:m + 1]
            hit = (neighbor_idx == col) & (col != -1)
            out_v = torch.where(hit, sv[:, m:m + 1], out_v)
            out_e = torch.where(hit, se[:, m:m + 1], out_e)
        return out_v, out_e
    nb = np.asarray(neighbor_idx)
    ex = np.asarray(excl_indices)
    sv = np.asarray(excl_scales_vdw, np.float32)
    se = np.asarray(excl_scales_elec, np.float32)
    out_v = np.ones(nb.shape, np.float32)
    out_e = np.ones(nb.shape, np.float32)
    for m in range(ex.shape[1]):
        col = ex[:, m:m + 1]
        hit = (nb == col) & (col != -1)
        out_v = np.where(hit, sv[:, m:m + 1], out_v)
        out_e = np.where(hit, se[:, m:m + 1], out_e)
    return out_v, out_e


def compute_nl_capacity(atom_bucket_size: int, box_size, cutoff: float, dr_threshold: float = 0.5,
                        safety_factor: float = 1.5, round_to: int = 256) -> int:
    """Static list capacity of a shape class (neighbor_list.py:451-530):
    ceil_to(round_to, safety_factor * (N / V) * 4/3 pi (cutoff + dr_threshold)^3)
    for a periodic box of volume V; ValueError for a non-positive volume."""
    import math

    vol = float(np.prod(np.asarray(box_size, np.float64)))
    if vol <= 0:
        raise ValueError(f"compute_nl_capacity: box_size must describe a positive-volume periodic box, "
                         f"got {box_size}")
    k_raw = safety_factor * (atom_bucket_size / vol) * (4.0 / 3.0) * math.pi * (cutoff + dr_threshold) ** 3
    return int(math.ceil(k_raw / round_to) * round_to)


@dataclasses.dataclass
class NeighborList:
    """A jax_md-style Dense neighbour list: ``idx`` (N, K) int32 device tensor
    with N after each row's last neighbour, built at ``reference_position``."""

    idx: object
    reference_position: object
    did_buffer_overflow: bool
    max_occupancy: int
    cutoff: float
    dr_threshold: float
    update_fn: object = None

    def update(self, positions, **kwargs) -> "NeighborList":
        return self.update_fn(positions, self, **kwargs)


class NeighborListFns:
    """``allocate`` / ``update`` of make_neighbor_list_fn."""

    def __init__(self, cutoff: float, dr_threshold: float, capacity_multiplier: float):
        self.cutoff = float(cutoff)
        self.dr_threshold = float(dr_threshold)
        self.capacity_multiplier = float(capacity_multiplier)

    def _build(self, positions, capacity: int | None):
        import ctypes

        import torch

        from prolix_amd import _native as N

        x = torch.as_tensor(positions, dtype=torch.float32, device="cuda").reshape(-1, 3).contiguous()
        n = x.shape[0]
        counts = torch.empty((n,), dtype=torch.int32, device="cuda")
        work = torch.empty((8 * ((n + 63) // 64),), dtype=torch.float32, device="cuda")
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        radius = self.cutoff + self.dr_threshold

        def run(k):
            idx = torch.empty((n, k), dtype=torch.int32, device="cuda")
            N.check(N.lib().prx_neighbor_list_dense(ctypes.c_void_p(x.data_ptr()), n, radius, k,
                                                    ctypes.c_void_p(idx.data_ptr()),
                                                    ctypes.c_void_p(counts.data_ptr()),
                                                    ctypes.c_void_p(work.data_ptr()), stream),
                    "prx_neighbor_list_dense")
            return idx

        if capacity is None:  # size from the occupancy, as jax_md's allocate does
            run(1)
            occ = int(counts.max().item()) if n else 0
            capacity = max(1, int(occ * self.capacity_multiplier))
        idx = run(capacity)
        occ = int(counts.max().item()) if n else 0
        return NeighborList(idx, x.clone(), occ > capacity, occ, self.cutoff, self.dr_threshold, self.update)

    def allocate(self, positions, extra_capacity: int = 0, **kwargs) -> NeighborList:
        nb = self._build(positions, None)
        if extra_capacity:
            nb = self._build(positions, nb.idx.shape[1] + int(extra_capacity))
        return nb

    def update(self, positions, neighbors: NeighborList, **kwargs) -> NeighborList:
        """Rebuild at the same capacity when an atom moved more than
        dr_threshold / 2 since the list was built; otherwise keep it."""
        import torch

        x = torch.as_tensor(positions, dtype=torch.float32, device="cuda").reshape(-1, 3)
        moved = torch.linalg.vector_norm(x - neighbors.reference_position, dim=1).max().item()
        if not moved > 0.5 * self.dr_threshold:
            return neighbors
        return self._build(x, neighbors.idx.shape[1])


def make_neighbor_list_fn(displacement_fn, box_size, cutoff: float, capacity_multiplier: float = 1.25,
                          disable_cell_list: bool = False, dr_threshold: float = 0.5) -> NeighborListFns:
    """Neighbour-list functions (neighbor_list.py:155-172): a Dense list of the
    pairs within cutoff + dr_threshold (0.5 A, the reference's skin), without
    self pairs, rebuilt on ``update`` when an atom moved more than half the
    skin. Free space only: ``box_size`` is not used for wrapping (the engine
    has no periodic boxes). The list is built on the GPU (tile boxes, then
    per-atom scans of the tiles in range); ``disable_cell_list`` has no
    effect."""
    return NeighborListFns(cutoff, dr_threshold, capacity_multiplier)
