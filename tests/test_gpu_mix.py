"""The 256-system batched mix at full size (425,157 atoms, 5.74e8 pairs per
step) in one packed engine -- the workload of the headline benchmark.

- a seeded 5% sample (13 systems) against the float64 C++ restatement, at
  the start positions and after 50 captured BAOAB steps: energy relative
  error <= 1e-6 and force RMS relative error <= 1e-5;
- two sampled systems take the same 50 steps alone in their own engines:
  positions, momenta and forces equal the mix's bit for bit;
- the engine's bonded branch on its own stream (on for this many atoms) and
  the bonded kernel in line give the same trajectory bit for bit.

Batched references: batched_energy.py:852-858, batched_simulate.py:218-382.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import cpu as C
from prolix_amd import synth
from tests.helpers import f32_round, rel_rms

pytestmark = pytest.mark.gpu

E_TOL = 1e-6
F_TOL = 1e-5
N_STEPS = 50


@pytest.fixture(scope="module")
def mix():
    print("[gpu] building the 256-system mix (spawned workers)", flush=True)
    prots = synth.build_many(synth.mix_jobs(256, 300, 5000, seed=2), start_method="spawn")
    systems = [f32_round(p.system_dict()) for p in prots]
    temps = synth.mix_temperatures(256, seed=2)
    rng = np.random.default_rng(2024)
    sample = sorted(rng.choice(256, size=13, replace=False).tolist())
    return systems, temps, sample


def _params(temps):
    from prolix_amd.engine import Langevin

    return [Langevin(1.0, float(temps[k]), 1.0, seed=1000 + k, safety_caps=True) for k in range(len(temps))]


def _check(systems, sample, e, f, offsets, positions, label):
    worst_e = worst_f = 0.0
    for k in sample:
        s = systems[k]
        x = positions[offsets[k]:offsets[k + 1]].astype(np.float64)
        eo, fo = C.CpuSystem(s).energy_forces(x, np.float64)
        e_rel = abs(e[k, -1] - eo[11]) / abs(eo[11])
        f_rel = rel_rms(f[offsets[k]:offsets[k + 1]], fo)
        print(f"[mix] {label} system {k} ({len(s['charges'])} atoms): energy rel {e_rel:.2e}, force rel {f_rel:.2e}",
              flush=True)
        assert e_rel <= E_TOL, (k, e_rel)
        assert f_rel <= F_TOL, (k, f_rel)
        worst_e, worst_f = max(worst_e, e_rel), max(worst_f, f_rel)
    return worst_e, worst_f


def test_full_mix_against_float64_and_alone(mix):
    from prolix_amd.engine import GBEngine

    systems, temps, sample = mix
    print("[gpu] packed engine for the mix", flush=True)
    eng = GBEngine(systems)
    assert eng.n_atoms == 425157
    off = eng.offsets
    x0 = np.concatenate([s["positions"] for s in systems]).astype(np.float32)
    e, f, _ = eng.energy_forces(x0, energies=True)
    _check(systems, sample, e.cpu().numpy(), f.cpu().numpy(), off, x0, "start")

    print(f"[gpu] {N_STEPS} captured steps of the mix", flush=True)
    eng.init(x0, _params(temps))
    eng.steps(N_STEPS)
    x, p, fs = (a.cpu().numpy() for a in eng.state())
    assert np.isfinite(x).all() and (eng.nonfinite_counts() == 0).all()
    e1, f1, _ = eng.energy_forces(x, energies=True)
    _check(systems, sample, e1.cpu().numpy(), f1.cpu().numpy(), off, x, f"after {N_STEPS} steps")
    # the force the step graph stored at the final positions is the force-only
    # evaluation there, bit for bit (the energy variant of the force pass may
    # round differently); caps leave it unchanged unless a force-cap event
    # happened in the system
    _, f0, _ = eng.energy_forces(x, energies=False)
    f0 = f0.cpu().numpy()
    for k in sample:
        if eng.warnings(k)[1] == 0:
            assert np.array_equal(fs[off[k]:off[k + 1]], f0[off[k]:off[k + 1]])

    print("[gpu] the same steps with the bonded kernel in line", flush=True)
    assert eng.n_atoms >= 100000  # the default fork threshold: the mix runs the branch
    eng.set_fork_min_atoms(10**9)
    eng.init(x0, _params(temps))
    eng.steps(N_STEPS)
    xi, pi_, fi = (a.cpu().numpy() for a in eng.state())
    eng.close()
    assert np.array_equal(xi, x) and np.array_equal(pi_, p) and np.array_equal(fi, fs)

    params = _params(temps)
    for k in sample[:2]:
        print(f"[gpu] system {k} alone", flush=True)
        one = GBEngine(systems[k])
        one.init(x0[off[k]:off[k + 1]], params[k])
        one.steps(N_STEPS)
        xa, pa, fa = (a.cpu().numpy() for a in one.state())
        one.close()
        sl = slice(off[k], off[k + 1])
        assert np.array_equal(xa, x[sl]) and np.array_equal(pa, p[sl]) and np.array_equal(fa, fs[sl])
