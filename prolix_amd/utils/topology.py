"""Bonded-graph exclusions: the atom pairs 1, 2 and 3 bonds apart.

Result contract of prolix's ``utils/topology.py:26-93``
(``find_bonded_exclusions``): unique pairs (i, j) with i < j, sorted, and each
pair in exactly one class, the one of its graph distance (a pair joined both
by a bond and by a longer path is a 1-2 pair only).

Method: graph distances up to 3 by frontier expansion over an edge list held
as sorted int64 pair keys (key = i * n + j). Distance-d pairs are the
endpoints of walks of length d that are not closer; a walk of length d is a
walk of length d - 1 extended by one edge, formed for all pairs at once with
a CSR neighbour table (no per-atom Python loop).
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np


class TopologyExclusions(NamedTuple):
    idx_12: np.ndarray  # (N12, 2) int32
    idx_13: np.ndarray  # (N13, 2) int32
    idx_14: np.ndarray  # (N14, 2) int32


def _keys_to_pairs(keys: np.ndarray, n: int) -> np.ndarray:
    """Sorted unique keys of ordered pairs -> (M, 2) int32 rows with i < j."""
    i, j = np.divmod(keys, n)
    up = i < j
    return np.stack([i[up], j[up]], axis=1).astype(np.int32) if up.any() else np.zeros((0, 2), np.int32)


def _extend(keys: np.ndarray, n: int, start: np.ndarray, nbr: np.ndarray) -> np.ndarray:
    """Every (a, c) with (a, b) in ``keys`` and c a neighbour of b (CSR
    ``start`` / ``nbr``), as sorted unique keys."""
    a, b = np.divmod(keys, n)
    deg = start[b + 1] - start[b]
    src = np.repeat(a, deg)
    # position of each output inside its source's neighbour run
    run = np.arange(deg.sum()) - np.repeat(np.cumsum(deg) - deg, deg)
    c = nbr[np.repeat(start[b], deg) + run]
    return np.unique(src * n + c)


def find_bonded_exclusions(bonds, n_atoms: int) -> TopologyExclusions:
    """1-2, 1-3 and 1-4 pairs of the bond graph (topology.py:26-93)."""
    n = int(n_atoms)
    b = np.asarray(bonds if bonds is not None else np.zeros((0, 2)), np.int64).reshape(-1, 2)
    b = b[b[:, 0] != b[:, 1]]  # a self-bond is no edge
    if b.shape[0] == 0 or n == 0:
        e = np.zeros((0, 2), np.int32)
        return TopologyExclusions(e, e, e)
    if b.min() < 0 or b.max() >= n:
        raise IndexError("bond index out of range")
    # both directions of every edge, unique, as sorted keys; CSR by source atom
    e1 = np.unique(np.concatenate([b[:, 0] * n + b[:, 1], b[:, 1] * n + b[:, 0]]))
    src, dst = np.divmod(e1, n)
    start = np.zeros(n + 1, np.int64)
    np.add.at(start, src + 1, 1)
    start = np.cumsum(start)
    nbr = dst  # already grouped by source (keys are sorted)
    self_keys = np.arange(n, dtype=np.int64) * (n + 1)
    w2 = _extend(e1, n, start, nbr)          # ends of 2-walks
    d2 = np.setdiff1d(w2, np.union1d(e1, self_keys), assume_unique=True)
    w3 = _extend(w2, n, start, nbr)          # ends of 3-walks
    closer = np.union1d(np.union1d(e1, d2), self_keys)
    d3 = np.setdiff1d(w3, closer, assume_unique=True)
    return TopologyExclusions(_keys_to_pairs(e1, n), _keys_to_pairs(d2, n), _keys_to_pairs(d3, n))
