"""The integrator-loop drop-ins against float64 restatements in the reference's
units (physics/simulate.py: ps, kcal/mol, masses in amu/418.4).

- nvt_langevin at gamma = 0 is velocity Verlet: every device step equals one
  float64 Verlet step in ps / (amu/418.4) units taken from the device's own
  previous state (physics/simulate.py:245-255, 421-431);
- run_minimization is steepest descent with a 4-point backtracking search
  (physics/simulate.py:141-196): each device iteration picks the candidate
  the float64 search picks, and lands where it lands;
- run_simulation is run_minimization then run_thermalization at the
  reference's schedule (physics/simulate.py:357-508).
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import gb_oracle as O
from tests.helpers import protein

pytestmark = pytest.mark.gpu


def _system():
    """The 138-atom peptide as a dict system: make_energy_fn takes no CMAP from
    a dict (PhysicsSystem.from_dict drops it), so neither does the oracle."""
    s = protein(138, 0)
    s.pop("cmap_torsions", None)
    return s


def test_nvt_langevin_zero_friction_is_float64_verlet_in_ps_units():
    from prolix_amd.physics import simulate as S
    from prolix_amd.physics.system import make_energy_fn

    print("[gpu] nvt_langevin gamma = 0, 138 atoms, 40 steps of 1 fs", flush=True)
    sysd = _system()
    fn = make_energy_fn(None, sysd, implicit_solvent=True, cutoff=0)
    x0 = S.run_minimization(fn, sysd["positions"], sd_steps=50).cpu().numpy().astype(np.float64)
    m = np.asarray(sysd["masses"], np.float64) / S.PS_MASS_TO_AMU  # amu/418.4 (physics/simulate.py:421-431)
    dt = 1e-3  # ps
    init_fn, apply_fn = S.nvt_langevin(fn, None, dt=dt, kT=S.BOLTZMANN_KCAL * 300.0, gamma=0.0, mass=m)
    st = init_fn(np.array([0, 3], np.uint32), x0)
    traj = [(st.positions.cpu().numpy().astype(np.float64), st.momentum.cpu().numpy().astype(np.float64))]
    for _ in range(40):
        st = apply_fn(st)
        traj.append((st.positions.cpu().numpy().astype(np.float64), st.momentum.cpu().numpy().astype(np.float64)))
    ff = lambda x: O.forces(sysd, x)

    worst_dx = worst_dp = 0.0
    for k in range(0, 40, 4):
        x, p = traj[k]
        p_half = p + 0.5 * dt * ff(x)          # kcal/mol/A = (amu/418.4) A/ps^2
        xv = x + dt * p_half / m[:, None]       # A
        pv = p_half + 0.5 * dt * ff(xv)
        xd, pd = traj[k + 1]
        worst_dx = max(worst_dx, float(np.max(np.abs(xd - xv))))
        worst_dp = max(worst_dp, float(np.sqrt(np.mean((pd - pv) ** 2) / np.mean(pv**2))))
    ulp = float(np.max(np.spacing(np.abs(x0).astype(np.float32))))
    # kinetic energy in the reference's units: sum p^2 / 2m = 3/2 N kT on average at start
    ke0 = float(np.sum(traj[0][1] ** 2 / (2.0 * m[:, None])))
    t0 = 2.0 * ke0 / (3.0 * len(m) * S.BOLTZMANN_KCAL)
    print(f"[dropin] per step max |dx| {worst_dx:.2e} A (ulp {ulp:.1e}), rel dp {worst_dp:.2e}, "
          f"initial T {t0:.1f} K", flush=True)
    assert worst_dx < 4.0 * ulp
    assert worst_dp < 1e-5
    assert 200.0 < t0 < 400.0  # Maxwell-Boltzmann draw at 300 K, read in ps units


def test_run_minimization_follows_float64_line_search():
    from prolix_amd.physics import simulate as S
    from prolix_amd.physics.system import make_energy_fn

    print("[gpu] run_minimization one iteration at a time vs float64, 138 atoms", flush=True)
    sysd = _system()
    fn = make_energy_fn(None, sysd, implicit_solvent=True, cutoff=0)
    efn = lambda x: O.total_energy(sysd, x)
    ffn = lambda x: O.forces(sysd, x)
    x = np.asarray(sysd["positions"], np.float64)
    chosen = []
    for it in range(12):
        xd = S.run_minimization(fn, x, sd_steps=1).cpu().numpy().astype(np.float64)
        xo, frac, es, e0 = O.sd_backtrack_step(efn, ffn, x)
        cand = sorted(es + [e0])
        # a near tie between the two lowest energies may fall either way in float32
        if cand[1] - cand[0] > 1e-5 * abs(cand[0]):
            assert np.max(np.abs(xd - xo)) < 1e-4, (it, frac)
            chosen.append(frac)
        x = xd
    print(f"[dropin] chosen fractions {chosen}", flush=True)
    assert len(chosen) >= 8
    assert efn(x) < efn(np.asarray(sysd["positions"], np.float64))


def test_run_simulation_is_minimisation_then_thermalisation():
    from prolix_amd.physics import simulate as S
    from prolix_amd.physics.system import make_energy_fn

    print("[gpu] run_simulation schedule, 138 atoms", flush=True)
    sysd = _system()
    key = np.array([0, 9], np.uint32)
    r = S.run_simulation(sysd, sysd["positions"], temperature=300.0, min_steps=50, therm_steps=120, key=key)
    fn = make_energy_fn(None, sysd, implicit_solvent=True)
    x = S.run_minimization(fn, sysd["positions"], steps=50)
    m = np.asarray(sysd["masses"], np.float64) / S.PS_MASS_TO_AMU
    r2 = S.run_thermalization(fn, x, temperature=300.0, steps=120, mass=m, key=key)
    assert fn._engine.step_count(0) == 120
    assert np.array_equal(r.cpu().numpy(), r2.cpu().numpy())
