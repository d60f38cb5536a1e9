"""Batched FIRE on the device against the float64 FIRE restatement, and the
steepest-descent minimiser's energy decrease."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import gb_oracle as O
from prolix_amd import synth
from tests.helpers import protein

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cap", (None, 10.0))
def test_fire_matches_oracle(cap):
    """30 FIRE steps (momentum form, per-system dt/alpha) from the reference's
    dynamic dt rule (simulate.py:642-656), with and without the capped apply."""
    from prolix_amd.engine import GBEngine
    from prolix_amd.minimize import fire

    s = protein(60, 4)
    m = np.asarray(s["masses"], np.float64)
    ff = lambda x: O.forces(s, x)
    f0 = ff(s["positions"])
    dt_start = float(np.clip(0.001 / (np.max(np.linalg.norm(f0, axis=1)) + 1e-8), 1e-7, 0.001))
    dt_max = min(dt_start * 10.0, 0.01)
    print(f"[gpu] FIRE 30 steps cap={cap}", flush=True)
    eng = GBEngine(s)
    st = fire(eng, s["positions"], 30, m, dt_start, dt_max, force_cap=cap)
    xg = st.positions.cpu().numpy().astype(np.float64)
    eng.close()
    xo = O.fire_minimize(ff, s["positions"], m, 30, dt_start, dt_max, force_cap=cap)[0]
    err = float(np.max(np.abs(xg - xo)))
    e_start = O.total_energy(s, s["positions"])
    e_dev, e_orc = O.total_energy(s, xg), O.total_energy(s, xo)
    print(f"[fire cap={cap}] max |dx| {err:.2e} A; E {e_start:.2f} -> device {e_dev:.4f}, oracle {e_orc:.4f}",
          flush=True)
    assert err < 1e-3
    assert e_dev < e_start
    # float32 positions: |F| ~ 30 kcal/mol/A times ~1e-6 A over 60 atoms
    assert abs(e_dev - e_orc) <= 1e-5 * abs(e_start)


def test_batched_fire_is_per_system():
    """Two systems minimised together follow the same path as each alone."""
    from prolix_amd.engine import GBEngine
    from prolix_amd.minimize import fire

    systems = [protein(n, 50 + k, radius=synth.compact_radius(n)) for k, n in enumerate((70, 130))]
    masses = [np.asarray(s["masses"], np.float64) for s in systems]
    print("[gpu] batched FIRE", flush=True)
    eng = GBEngine(systems)
    st = fire(eng, [s["positions"] for s in systems], 40, np.concatenate(masses), 2e-5, 2e-4, force_cap=100.0)
    xb = eng.unpack(st.positions)
    eng.close()
    for k, s in enumerate(systems):
        e1 = GBEngine(s)
        st1 = fire(e1, s["positions"], 40, masses[k], 2e-5, 2e-4, force_cap=100.0)
        x1 = st1.positions.cpu().numpy()
        e1.close()
        assert np.max(np.abs(x1 - xb[k])) < 1e-5, k


@pytest.mark.parametrize("k", (0.0, 500.0))
def test_capped_descent_matches_oracle(k):
    """20 steps of the batched minimiser's first stage (per-atom capped steepest
    descent, batched_simulate.py:746-787), with and without a heavy-atom restraint."""
    from prolix_amd.engine import GBEngine
    from prolix_amd.minimize import capped_descent

    s = protein(60, 4)
    heavy = (np.asarray(s["masses"]) > 2.0).astype(np.float64)
    print(f"[gpu] capped descent 20 steps k={k}", flush=True)
    eng = GBEngine(s)
    xg = capped_descent(eng, s["positions"], 20, k_restraint=k, restrained=heavy).cpu().numpy().astype(np.float64)
    eng.close()
    xo = O.capped_descent(lambda x: O.forces(s, x), s["positions"], 20, k_restraint=k, restrained=heavy)
    err = float(np.max(np.abs(xg - xo)))
    print(f"[capped k={k}] max |dx| {err:.2e} A", flush=True)
    # float32 positions (~2e-6 A at 20 A) times k = 500 put ~1e-3 kcal/mol/A
    # into the restraint force each step; without it the bound is 1e-4 A
    assert err < (5e-4 if k > 0 else 1e-4)
    assert O.total_energy(s, xg) < O.total_energy(s, s["positions"])


def test_restrained_fire_stage_matches_oracle():
    """One restrained FIRE sub-stage of the batched minimiser
    (batched_simulate.py:813-874): restraint 500 kcal/mol/A^2 on every third
    atom toward the start, 0.1 A step cap, best snapshot."""
    from prolix_amd.engine import GBEngine
    from prolix_amd.minimize import restrained_fire

    s = protein(60, 4)
    m = np.asarray(s["masses"], np.float64)
    x0 = np.asarray(s["positions"], np.float64)
    sel = (np.arange(60) % 3 == 0).astype(np.float64)
    ff = lambda x: O.forces(s, x) - 500.0 * (x - x0) * sel[:, None]
    print("[gpu] restrained FIRE stage, 30 steps", flush=True)
    eng = GBEngine(s)
    xg, rg = restrained_fire(eng, x0, m, 30, k_restraint=500.0, selection=sel)
    xg, rg = xg.cpu().numpy().astype(np.float64), float(rg.cpu().numpy()[0])
    eng.close()
    xo, ro = O.fire_stage_best(ff, x0, m, 30)
    err = float(np.max(np.abs(xg - xo)))
    print(f"[fire stage] max |dx| {err:.2e} A, best rms device {rg:.5f} oracle {ro:.5f}", flush=True)
    assert err < 1e-3
    assert abs(rg - ro) <= 1e-4 * ro


def test_lbfgs_matches_oracle():
    """The L-BFGS polish: 10 iterations of the same method in float64."""
    from prolix_amd.engine import GBEngine
    from prolix_amd.minimize import lbfgs

    s = protein(60, 4)
    print("[gpu] L-BFGS 10 iterations", flush=True)
    eng = GBEngine(s)
    xg = lbfgs(eng, s["positions"], steps=10).cpu().numpy().astype(np.float64)
    eng.close()
    xo, tr = O.lbfgs_minimize(lambda x: O.total_energy(s, x), lambda x: -O.forces(s, x), s["positions"], 10)
    e0, eg, eo = O.total_energy(s, s["positions"]), O.total_energy(s, xg), O.total_energy(s, xo)
    err = float(np.max(np.abs(xg - xo)))
    print(f"[lbfgs] max |dx| {err:.2e} A; E {e0:.3f} -> device {eg:.4f}, oracle {eo:.4f}", flush=True)
    assert np.all(np.diff(tr) <= 0) and eo < e0
    assert err < 2e-3
    assert abs(eg - eo) <= 1e-5 * abs(e0)


def test_batched_minimize_stages_per_system():
    """The six stages on two systems together; each equals its run alone to
    float32 rounding, and every system ends below its starting energy."""
    from prolix_amd.engine import GBEngine
    from prolix_amd.minimize import batched_minimize

    systems = [protein(n, 60 + k, radius=synth.compact_radius(n)) for k, n in enumerate((80, 150))]
    masses = np.concatenate([np.asarray(s["masses"], np.float64) for s in systems])
    kw = dict(max_steps=40, lbfgs_steps=15, fire_stage_steps=(20, 20, 20, 20))
    print("[gpu] batched minimiser, 2 systems", flush=True)
    eng = GBEngine(systems)
    x, conv, rms = batched_minimize(eng, [s["positions"] for s in systems], masses, **kw)
    xb = eng.unpack(x)
    e_start = eng.energy_forces([s["positions"] for s in systems])[0].cpu().numpy()[:, -1]
    e_end = eng.energy_forces(x)[0].cpu().numpy()[:, -1]
    eng.close()
    assert np.all(e_end < e_start) and np.isfinite(xb[0]).all() and conv.shape == (2,)
    off = 0
    for k, s in enumerate(systems):
        n = len(s["charges"])
        one = GBEngine(s)
        xa, _, ra = batched_minimize(one, s["positions"], masses[off:off + n], **kw)
        one.close()
        off += n
        assert np.max(np.abs(xa.cpu().numpy() - xb[k])) < 1e-4


def test_masses_as_device_tensor():
    """The FIRE stages take the masses as an array or as a tensor on the
    device, with the same result."""
    import torch

    from prolix_amd.engine import GBEngine
    from prolix_amd.minimize import fire, restrained_fire

    s = protein(60, 4)
    m = np.asarray(s["masses"], np.float64)
    mt = torch.as_tensor(m, device="cuda")
    print("[gpu] FIRE with masses as array and as device tensor", flush=True)
    eng = GBEngine(s)
    a = fire(eng, s["positions"], 10, m, 1e-4, 1e-3).positions.cpu().numpy()
    b = fire(eng, s["positions"], 10, mt, 1e-4, 1e-3).positions.cpu().numpy()
    ra, _ = restrained_fire(eng, s["positions"], m, 10, k_restraint=25.0)
    rb, _ = restrained_fire(eng, s["positions"], mt, 10, k_restraint=25.0)
    eng.close()
    assert np.array_equal(a, b)
    assert np.array_equal(ra.cpu().numpy(), rb.cpu().numpy())
