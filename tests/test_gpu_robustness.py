"""Integrator state, resume and failure handling on the device.

- an energy evaluation at other positions between steps leaves the
  integrator's trajectory unchanged bit for bit;
- nvt_langevin's apply_fn is a function of its state argument;
- a saved state resumes its Langevin trajectory bit for bit (step counter =
  Philox counter, batched_simulate.py:1300-1305);
- a system that turns non-finite is flagged on the device, evicted by the
  serving loop with an error result, and the other systems are unchanged;
- a non-finite starting energy raises (simulate.py:571-582);
- the bonded kernel on its own stream (a parallel graph branch) gives the same
  trajectory as in line.
NaN arithmetic on the GPU is ordinary floating point (no fault); every
index the kernels form stays in range for any coordinate value.
"""

from __future__ import annotations

import numpy as np
import pytest

from prolix_amd import synth
from tests.helpers import protein

pytestmark = pytest.mark.gpu


def _systems():
    return [protein(int(n), 60 + k, radius=synth.compact_radius(int(n))) for k, n in enumerate((300, 700, 1100))]


def _params(n, caps=True):
    from prolix_amd.engine import Langevin

    return [Langevin(1.0, 300.0 + 10 * k, 1.0, seed=21 + k, safety_caps=caps) for k in range(n)]


def _run(eng, systems, steps, between=None):
    eng.init([s["positions"] for s in systems], _params(len(systems)))
    eng.steps(steps // 2)
    if between is not None:
        between(eng)
    eng.steps(steps - steps // 2)
    return [a.cpu().numpy() for a in eng.state()]


def test_energy_call_between_steps_leaves_the_trajectory_unchanged():
    from prolix_amd.engine import GBEngine

    print("[gpu] energy_forces at other positions between steps, 3 systems", flush=True)
    systems = _systems()
    eng = GBEngine(systems)
    ref = _run(eng, systems, 20)
    rng = np.random.default_rng(0)
    other = [s["positions"] + rng.normal(scale=0.3, size=s["positions"].shape) for s in systems]
    got = _run(eng, systems, 20, between=lambda e: e.energy_forces(other, energies=True, born_radii=True))
    eng.close()
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)


def test_apply_fn_is_a_function_of_its_state():
    from prolix_amd.physics import simulate as S
    from prolix_amd.physics.system import make_energy_fn

    print("[gpu] nvt_langevin apply_fn on older states, 138 atoms", flush=True)
    p = synth.peptide_138(0)
    sysd = p.system_dict()
    fn = make_energy_fn(None, sysd, implicit_solvent=True, cutoff=0)
    m = np.asarray(sysd["masses"], np.float64) / S.PS_MASS_TO_AMU
    init_fn, apply_fn = S.nvt_langevin(fn, None, dt=1e-3, kT=S.BOLTZMANN_KCAL * 300.0, gamma=1.0, mass=m)
    # from relaxed positions: the raw synthetic start blows up within 20 steps
    # without the batched step's caps
    x0 = S.run_minimization(fn, sysd["positions"], sd_steps=50)
    s0 = init_fn(np.array([0, 7], np.uint32), x0)
    s1 = apply_fn(s0, n_steps=10)
    x1 = s1.positions.cpu().numpy()
    s2 = apply_fn(s1, n_steps=10)
    x2, p2 = s2.positions.cpu().numpy(), s2.momentum.cpu().numpy()
    assert np.isfinite(x2).all() and np.isfinite(p2).all()
    fn(sysd["positions"] + 0.1)  # an energy call at other positions
    s1b = apply_fn(s0, n_steps=10)  # an older state: replayed, not continued
    assert s1b.step == 10 and np.array_equal(s1b.positions.cpu().numpy(), x1)
    s2b = apply_fn(s1, n_steps=10)
    assert s2b.step == 20
    assert np.array_equal(s2b.positions.cpu().numpy(), x2) and np.array_equal(s2b.momentum.cpu().numpy(), p2)
    s2c = apply_fn(s1b, n_steps=5)
    s2c = apply_fn(s2c, n_steps=5)  # 10 steps in two calls from an equal state
    assert np.array_equal(s2c.positions.cpu().numpy(), x2)


def test_saved_state_resumes_the_trajectory_bit_for_bit(tmp_path):
    from prolix_amd import io
    from prolix_amd.engine import GBEngine

    print("[gpu] 100 steps vs 50 + save/load + 50, 3 systems", flush=True)
    systems = _systems()
    masses = np.concatenate([np.asarray(s["masses"], np.float64) for s in systems])
    eng = GBEngine(systems)
    eng.init([s["positions"] for s in systems], _params(3))
    eng.steps(100)
    x_ref, p_ref, f_ref = (a.cpu().numpy() for a in eng.state())
    eng.init([s["positions"] for s in systems], _params(3))
    eng.steps(50)
    states = io.engine_states(eng, masses, dt_fs=1.0)
    eng.close()
    path = tmp_path / "resume.prx"
    with io.TrajectoryWriter(str(path)) as w:
        w.write(states)
    back = list(io.read_trajectory(str(path)))
    assert [int(st.step) for st in back] == [50, 50, 50]
    eng2 = GBEngine(systems)  # a fresh engine, as after a restart
    io.restore_engine(eng2, back, _params(3))
    eng2.steps(50)
    x, p, f = (a.cpu().numpy() for a in eng2.state())
    counts = [eng2.step_count(k) for k in range(3)]
    eng2.close()
    assert counts == [100, 100, 100]
    assert np.array_equal(x, x_ref) and np.array_equal(p, p_ref) and np.array_equal(f, f_ref)


def test_nonfinite_system_is_flagged_and_evicted_others_unchanged():
    from prolix_amd.engine import GBPool, Langevin
    from prolix_amd.serving import Request, serve

    print("[gpu] pool with one NaN-position system", flush=True)
    systems = _systems()
    bad = dict(systems[1])
    bad["positions"] = systems[1]["positions"].copy()
    bad["positions"][5] = np.nan
    pool = GBPool(max_atoms=4000, max_systems=8)
    slots = [pool.admit(s, Langevin(1.0, 300.0, 1.0, seed=3 + k, safety_caps=True))
             for k, s in enumerate([systems[0], bad, systems[2]])]
    pool.steps(5)
    counts = pool.nonfinite_counts()
    pool.close()
    assert counts[slots[0]] == 0 and counts[slots[2]] == 0 and counts[slots[1]] > 0

    def trace(with_bad):
        reqs = [Request(systems[0], 300, 300.0, 11, 0.0, rid=0), Request(systems[2], 500, 320.0, 12, 0.0, rid=2)]
        if with_bad:
            reqs.append(Request(bad, 400, 310.0, 13, 0.0, rid=1))
        return serve(reqs, hbm_budget_bytes=1e10, chunk_steps=100)

    print("[gpu] serving with and without the NaN request", flush=True)
    clean, st_clean = trace(False)
    mixed, st_mixed = trace(True)
    assert st_clean["failed"] == 0 and st_mixed["failed"] == 1
    by = {r.rid: r for r in mixed}
    assert by[1].error is not None and by[1].n_steps < 400  # evicted at its first check
    for r in clean:
        m = by[r.rid]
        assert m.error is None and m.n_steps == r.n_steps
        assert np.array_equal(m.positions, r.positions) and np.array_equal(m.momenta, r.momenta)


def test_engine_health_check_raises_for_the_failed_system():
    from prolix_amd.checks import SimulationInstability, check_engine_health, check_initial_energy
    from prolix_amd.engine import GBEngine

    print("[gpu] health check and initial-energy check", flush=True)
    systems = _systems()[:2]
    x = [s["positions"].copy() for s in systems]
    x[1][3] = np.inf
    eng = GBEngine(systems)
    with pytest.raises(RuntimeError, match="non-finite"):
        check_initial_energy(eng, x)
    tot = check_initial_energy(eng, [s["positions"] for s in systems])
    assert np.isfinite(tot).all()
    eng.init(x, _params(2))
    eng.steps(3)
    with pytest.raises(SimulationInstability) as ei:
        check_engine_health(eng)
    assert ei.value.systems == [1]
    eng.close()


def test_bonded_branch_equals_in_line():
    from prolix_amd.engine import GBEngine

    print("[gpu] bonded forked vs in line, 3 systems", flush=True)
    systems = _systems()
    eng = GBEngine(systems)
    eng.set_fork_min_atoms(10**9)
    inline = _run(eng, systems, 30)
    eng.set_fork_min_atoms(0)
    forked = _run(eng, systems, 30)
    e_f = eng.energy_forces([s["positions"] for s in systems])[0].cpu().numpy()
    eng.set_fork_min_atoms(10**9)
    e_i = eng.energy_forces([s["positions"] for s in systems])[0].cpu().numpy()
    eng.close()
    for a, b in zip(inline, forked):
        assert np.array_equal(a, b)
    assert np.array_equal(e_f, e_i)
