"""Bonded-graph exclusions: 1-2, 1-3 and 1-4 pairs from a bond list.

Mirrors prolix's ``utils/topology.py:26-93`` (``find_bonded_exclusions``):
pairs are unique, stored (i, j) with i < j, and a closer relation wins
(1-3 pairs that are also 1-2 pairs are dropped, 1-4 pairs that are 1-2 or
1-3 pairs are dropped).
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np


class TopologyExclusions(NamedTuple):
    idx_12: np.ndarray  # (N12, 2) int32
    idx_13: np.ndarray  # (N13, 2) int32
    idx_14: np.ndarray  # (N14, 2) int32


def _pairs(s) -> np.ndarray:
    if not s:
        return np.zeros((0, 2), dtype=np.int32)
    return np.array(sorted(s), dtype=np.int32)


def find_bonded_exclusions(bonds, n_atoms: int) -> TopologyExclusions:
    """All 1-2, 1-3 and 1-4 pairs of the bond graph (topology.py:26-93)."""
    if bonds is None or len(bonds) == 0:
        e = np.zeros((0, 2), dtype=np.int32)
        return TopologyExclusions(e, e, e)
    bonds = np.asarray(bonds, dtype=np.int64)
    adj = [set() for _ in range(n_atoms)]
    for a, b in bonds:
        if a == b:
            continue
        adj[a].add(int(b))
        adj[b].add(int(a))
    s12, s13, s14 = set(), set(), set()
    for i in range(n_atoms):
        for j in adj[i]:
            if j > i:
                s12.add((i, j))
            for k in adj[j]:
                if k == i:
                    continue
                if k > i:
                    s13.add((i, k))
                for l in adj[k]:
                    if l in (j, i):
                        continue
                    if l > i:
                        s14.add((i, l))
    s13 -= s12
    s14 -= s12
    s14 -= s13
    return TopologyExclusions(_pairs(s12), _pairs(s13), _pairs(s14))
