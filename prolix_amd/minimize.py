"""Steepest descent with a 4-point backtracking search, batched on the device.

Restates prolix's ``run_minimization`` (physics/simulate.py:141-196): per
iteration one energy+force evaluation at r, then energies at
r + {1, 0.5, 0.25, 0.1} * dt * F with dt = min(max_disp / max|F|, dt_max);
the lowest of those candidates replaces r if it is below E(r). Every system
of the engine takes its own step; all choices are made with device tensor
ops, so the loop never waits on the host.
"""

from __future__ import annotations

import numpy as np

FRACTIONS = (1.0, 0.5, 0.25, 0.1)


def minimize(engine, positions, steps: int = 200, dt_max: float = 4e-3, max_displacement: float = 0.1):
    """Returns the relaxed packed positions (device float32 tensor)."""
    t = engine.torch
    x = engine._pos_tensor(positions).clone()
    sid = t.as_tensor(np.repeat(np.arange(engine.n_systems), engine.sizes), device=x.device)
    total = 11  # index of the total energy
    for _ in range(int(steps)):
        e, f, _ = engine.energy_forces(x, energies=True)
        e0 = e[:, total]
        fmag = t.linalg.vector_norm(f, dim=1)
        fmax = t.zeros(engine.n_systems, dtype=f.dtype, device=f.device).scatter_reduce(
            0, sid, fmag, reduce="amax", include_self=True) + 1e-8
        dt = t.clamp(max_displacement / fmax, max=dt_max)
        best_x = x
        best_e = e0
        # The reference takes r4 (0.1) as the starting candidate and lets
        # r3, r2, r1 replace it when strictly lower (physics/simulate.py:186-191).
        cands = []
        for frac in FRACTIONS:
            xc = x + (frac * dt)[sid, None].to(x.dtype) * f
            ec, _, _ = engine.energy_forces(xc, energies=True)
            cands.append((xc, ec[:, total]))
        cx, ce = cands[3]
        for xc, ec in (cands[2], cands[1], cands[0]):
            better = ec < ce
            cx = t.where(better[sid, None], xc, cx)
            ce = t.where(better, ec, ce)
        accept = ce < best_e
        x = t.where(accept[sid, None], cx, best_x)
    return x
