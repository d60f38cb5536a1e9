"""Failure detection of the simulation loops.

- ``check_initial_energy``: a non-finite starting energy raises before any
  minimisation or dynamics (simulate.py:571-582).
- ``check_position_sanity``: a system whose largest coordinate magnitude
  exceeds a bound raises, naming every such system (batched_simulate.py:158-205).
- ``check_engine_health``: the device keeps, per system, a count of atom-steps
  that left a non-finite position or momentum (the NaN/Inf check the
  reference makes on each saved epoch, simulate.py:1050-1071); this reads the
  counts and raises for any system with one.
"""

from __future__ import annotations

import numpy as np

_TOTAL = 11  # index of the total energy in an engine energy row


class SimulationInstability(RuntimeError):
    """A system became non-finite or left the sane coordinate range."""

    def __init__(self, msg: str, systems):
        super().__init__(msg)
        self.systems = list(systems)


def check_initial_energy(engine, positions) -> np.ndarray:
    """Total energy of every system at ``positions``; raises RuntimeError if one
    is non-finite (severe clashes or unparameterised atoms in the reference's
    wording). Returns the (n_systems,) totals."""
    e, _, _ = engine.energy_forces(positions, energies=True)
    tot = e[:, _TOTAL].cpu().numpy()
    bad = np.flatnonzero(~np.isfinite(tot))
    if bad.size:
        raise RuntimeError(
            f"Initial energy is non-finite for system(s) {bad.tolist()} "
            f"({', '.join(f'{tot[k]:.2f}' for k in bad)} kcal/mol). This typically indicates severe steric "
            "clashes or unparameterised atoms (sigma = 0).")
    return tot


def check_position_sanity(positions, sizes, max_magnitude: float, stage: str, system_names=None) -> None:
    """Raise ``SimulationInstability`` if a system's largest |coordinate| exceeds
    ``max_magnitude`` (A) or is non-finite. ``positions`` is packed (N, 3) (a
    tensor or array), ``sizes`` the atoms per system in packed order."""
    x = positions.detach().cpu().numpy() if hasattr(positions, "detach") else np.asarray(positions)
    x = x.reshape(-1, 3)
    failed = []
    off = 0
    for k, n in enumerate(np.asarray(sizes, np.int64)):
        part = x[off:off + n]
        off += n
        m = float(np.max(np.abs(part))) if n else 0.0
        if not np.isfinite(part).all() or m > max_magnitude:
            name = system_names[k] if system_names is not None and k < len(system_names) else f"sys_{k}"
            failed.append((k, name, m))
    if failed:
        details = ", ".join(f"{name}={m:.1f} A" for _, name, m in failed)
        raise SimulationInstability(
            f"MD explosion detected at {stage} stage: {len(failed)}/{len(sizes)} systems exceeded "
            f"{max_magnitude} A or are non-finite: {details}", [k for k, _, _ in failed])


def check_engine_health(engine, stage: str = "production") -> None:
    """Raise ``SimulationInstability`` naming every system of ``engine`` whose
    positions or momenta turned non-finite since its integrator init."""
    counts = engine.nonfinite_counts()
    bad = np.flatnonzero(counts > 0)
    if bad.size:
        raise SimulationInstability(
            f"Simulation instability at {stage}: non-finite positions or momenta in system(s) {bad.tolist()} "
            f"({counts[bad].tolist()} atom-steps). Reduce the time step or minimise further.", bad.tolist())
