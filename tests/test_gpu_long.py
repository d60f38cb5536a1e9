"""Parity after each configuration's full dynamics.

The 2,500-atom globule runs 10,000 BAOAB steps (no cutoff). The 20,000-atom
globule runs 20,000 steps with a 20 A cutoff in the neighbour-list GB form.
Both use dt 1 fs, 300 K and gamma 1/ps, after 2,000 steps of steepest descent.
The batched step's safety caps are on: the recipe's opposite-charge light
atoms can collapse.

At the final float32 positions, the device's energy and forces are compared
with the float64 C++ restatement on identical inputs. For the cutoff system
the restatement classifies pairs with the kernels' float32 r^2 test. Atoms
whose force reaches the cap are left out of the force comparison: the stored
force there is the capped one.
"""

from __future__ import annotations

import numpy as np
import pytest

from prolix_amd import synth
from tests.helpers import protein

pytestmark = pytest.mark.gpu

E_TOL = 1e-6
F_TOL = 1e-5


def _run(sysd, steps, **kw):
    import torch

    from prolix_amd.engine import GBEngine, Langevin
    from prolix_amd.minimize import minimize

    eng = GBEngine(sysd, **kw)
    x = minimize(eng, sysd["positions"], steps=2000)
    eng.init(x, Langevin(1.0, 300.0, 1.0, seed=21, safety_caps=True))
    eng.steps(steps)
    torch.cuda.synchronize()
    x, p, f = (a.cpu().numpy() for a in eng.state())
    e = eng.energy_forces(eng.state()[0])[0].cpu().numpy()[0]
    ke = float(eng.kinetic().cpu().numpy()[0])
    assert eng.step_count(0) == steps
    eng.close()
    return x, f, e, ke


def _check(name, sysd, x, f, e, ke, **okw):
    from oracle import cpu as C

    eo, fo = C.CpuSystem(sysd, **okw).energy_forces(x.astype(np.float64))
    n = len(sysd["charges"])
    temp = 2.0 * ke / (3.0 * n * 0.0019872041)
    keep = np.linalg.norm(fo, axis=1) < 9e3
    rel_e = abs(e[-1] - eo[-1]) / max(abs(eo[-1]), 0.2 * float(np.sum(np.abs(eo[:-1]))))
    err = f[keep].astype(np.float64) - fo[keep]
    rel_f = float(np.sqrt((err ** 2).sum() / (fo[keep] ** 2).sum()))
    print(f"[{name}] T {temp:.0f} K, total rel {rel_e:.2e}, force rel-rms {rel_f:.2e} "
          f"({n - int(keep.sum())} capped atoms left out)", flush=True)
    assert np.isfinite(x).all() and np.isfinite(f).all()
    assert rel_e < E_TOL
    assert rel_f < F_TOL
    assert keep.mean() > 0.99


def test_globule_2500_after_10000_steps():
    sysd = protein(2500, 1, radius=synth.compact_radius(2500))
    print("[gpu] globule 2500, 10000 steps", flush=True)
    x, f, e, ke = _run(sysd, 10000)
    _check("globule_2500_10k", sysd, x, f, e, ke)


def test_gb_cutoff_20k_after_20000_steps():
    sysd = protein(20000, 7, radius=synth.compact_radius(20000))
    print("[gpu] 20k atoms, 20 A cutoff, 20000 steps", flush=True)
    x, f, e, ke = _run(sysd, 20000, cutoff=20.0, gb_cutoff=True)
    _check("gbcut_20k_20k", sysd, x, f, e, ke, cutoff=20.0, gb_cutoff=True, f32_cutoff=True)
