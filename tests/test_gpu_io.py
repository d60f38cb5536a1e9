"""Engine states recorded through the trajectory writer (prolix/simulate.py:148-203)."""

from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import protein

pytestmark = pytest.mark.gpu


def test_engine_states_round_trip_through_a_trajectory(tmp_path):
    from prolix_amd import io
    from prolix_amd.engine import GBEngine, Langevin

    print("[gpu] engine_states, 2 systems of 60 / 95 atoms", flush=True)
    systems = [protein(60, 3), protein(95, 4)]
    eng = GBEngine(systems)
    masses = np.concatenate([np.asarray(s["masses"], np.float64) for s in systems])
    eng.init([s["positions"] for s in systems], [Langevin(0.5, 300.0, 1.0, seed=5), Langevin(0.5, 310.0, 1.0, seed=6)])
    eng.steps(40)
    states = io.engine_states(eng, masses, dt_fs=0.5)
    x, p, f = (a.cpu().numpy() for a in eng.state())
    e = eng.energy_forces(eng.state()[0])[0].cpu().numpy()
    ke = eng.kinetic().cpu().numpy()
    eng.close()

    path = tmp_path / "engine.prx"
    with io.TrajectoryWriter(str(path)) as w:
        w.write(states)
    back = list(io.read_trajectory(str(path)))
    assert len(back) == 2
    off = np.cumsum([0] + [len(s["charges"]) for s in systems])
    for k, st in enumerate(back):
        sl = slice(off[k], off[k + 1])
        assert st.step == 40 and st.time_ns == pytest.approx(40 * 0.5e-6, rel=1e-12)
        assert np.array_equal(st.positions, x[sl]) and np.array_equal(st.forces, f[sl])
        assert np.allclose(st.velocities * masses[sl, None], p[sl], rtol=1e-6, atol=1e-7)
        assert st.potential_energy == e[k, -1]
        # 1 amu A^2 / tau^2 = 1 kcal/mol in AKMA units.
        ke_v = 0.5 * float((masses[sl, None] * st.velocities.astype(np.float64) ** 2).sum())
        assert st.kinetic_energy == ke[k] and ke_v == pytest.approx(ke[k], rel=1e-5)
