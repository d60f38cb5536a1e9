"""Drop-in API, noise stream and serving pool on the device."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import gb_oracle as O
from oracle import philox
from prolix_amd import synth
from tests.helpers import f32_round, protein

pytestmark = pytest.mark.gpu

KB = 0.0019872041


def test_make_energy_fn_dropin_matches_oracle():
    """make_energy_fn(displacement_fn, system, implicit_solvent=True, exclusion_spec=spec, cutoff=0)."""
    from prolix_amd.physics.neighbor_list import ExclusionSpec
    from prolix_amd.physics.system import make_energy_fn

    p = synth.peptide_138(0)
    sysd = f32_round(p.as_dict())
    sysd.pop("cmap_torsions")  # dict systems carry no CMAP (PhysicsSystem.from_dict)
    spec = ExclusionSpec.from_bonds(sysd["bonds"], p.n_atoms)
    print("[gpu] make_energy_fn", flush=True)
    fn = make_energy_fn(None, sysd, implicit_solvent=True, exclusion_spec=spec, cutoff=0)
    e = fn(sysd["positions"], soft_core_lambda=1.0)  # unknown kwargs ignored (system.py:229)
    ref = dict(sysd)
    ref["excl_indices"], ref["excl_scales_vdw"], ref["excl_scales_elec"] = \
        __import__("prolix_amd.physics.neighbor_list", fromlist=["x"]).map_exclusions_to_dense_padded(spec)
    terms = O.energy_terms(ref, sysd["positions"])
    assert e == pytest.approx(terms["total"], rel=1e-6)
    f = fn.forces(sysd["positions"]).cpu().numpy()
    fo = O.forces(ref, sysd["positions"])
    assert np.sqrt(np.mean((f - fo) ** 2) / np.mean(fo**2)) < 1e-5
    d = make_energy_fn(None, sysd, implicit_solvent=True, exclusion_spec=spec, cutoff=0, return_decomposed=True)
    e_gb, e_direct, born = d["electrostatics"](sysd["positions"])
    assert e_gb == pytest.approx(terms["gb"], rel=1e-6)
    assert e_direct == pytest.approx(terms["coulomb"], rel=1e-6)
    assert np.allclose(born.cpu().numpy(), terms["_born_radii"], rtol=1e-5)
    # default cutoff is 9 A for LJ/Coulomb (system.py:134)
    fn9 = make_energy_fn(None, sysd, implicit_solvent=True, exclusion_spec=spec)
    assert fn9(sysd["positions"]) == pytest.approx(O.energy_terms(ref, sysd["positions"], cutoff=9.0)["total"],
                                                   rel=1e-6)
    with pytest.raises(ValueError):
        make_energy_fn(None, {"charges": np.zeros(0), "sigmas": np.zeros(0), "epsilons": np.zeros(0)})


def test_run_simulation_and_nvt_langevin():
    from prolix_amd.physics import simulate as S
    from prolix_amd.physics.neighbor_list import ExclusionSpec
    from prolix_amd.physics.system import make_energy_fn

    p = synth.peptide_138(0)
    sysd = p.system_dict()
    spec = ExclusionSpec.from_bonds(sysd["bonds"], p.n_atoms)
    print("[gpu] run_simulation", flush=True)
    r = S.run_simulation(sysd, sysd["positions"], temperature=300.0, min_steps=50, therm_steps=200,
                         key=np.array([0, 7]), exclusion_spec=spec)
    assert r.shape == (138, 3) and bool(np.isfinite(r.cpu().numpy()).all())
    fn = make_energy_fn(None, sysd, implicit_solvent=True, exclusion_spec=spec, cutoff=0)
    init_fn, apply_fn = S.nvt_langevin(fn, None, dt=1e-3, kT=KB * 300.0, gamma=1.0,
                                       mass=np.asarray(sysd["masses"]) / 418.4)
    st = init_fn(np.array([0, 1]), r)
    st = apply_fn(st, n_steps=20)
    assert bool(np.isfinite(st.position.cpu().numpy()).all())


def test_initial_momenta_follow_philox_stream():
    """Device draws equal the oracle's Philox4x32-10 + Box-Muller stream."""
    from prolix_amd.engine import GBEngine, Langevin

    s = protein(200, 3)
    eng = GBEngine(s)
    eng.init(s["positions"], Langevin(1.0, 310.0, 1.0, seed=123456789))
    p = eng.state()[1].cpu().numpy().astype(np.float64)
    eng.close()
    ref = philox.initial_momenta(123456789, s["masses"], KB * 310.0)
    assert np.max(np.abs(p - ref)) < 2e-5 * np.max(np.abs(ref))


def _pool_run(systems_in_order, schedule, params):
    """schedule: list of ('admit', k) | ('steps', n) | ('read', k)."""
    from prolix_amd.engine import GBPool

    pool = GBPool(max_atoms=sum(len(s["charges"]) for s in systems_in_order) + 10, max_systems=16)
    slots = {}
    out = {}
    for op, arg in schedule:
        if op == "admit":
            slots[arg] = pool.admit(systems_in_order[arg], params[arg])
        elif op == "steps":
            pool.steps(arg)
        elif op == "read":
            out[arg] = pool.read(slots[arg])
            pool.release(slots[arg])
    pool.close()
    return out


def test_pool_result_is_the_same_alone_and_in_the_mix():
    from prolix_amd.engine import Langevin

    systems = [protein(n, 60 + k, radius=synth.compact_radius(n)) for k, n in enumerate((150, 420, 97, 260))]
    params = [Langevin(1.0, 290.0 + 10 * k, 1.0, seed=900 + k, safety_caps=True) for k in range(4)]
    print("[gpu] pool alone", flush=True)
    alone = _pool_run(systems, [("admit", 0), ("steps", 150), ("read", 0)], params)
    print("[gpu] pool mix", flush=True)
    mixed = _pool_run(systems, [("admit", 1), ("admit", 2), ("steps", 40), ("admit", 0), ("steps", 30),
                                ("read", 2), ("admit", 3), ("steps", 120), ("read", 0), ("read", 1)], params)
    xa, pa, fa, na = alone[0]
    xm, pm, fm, nm = mixed[0]
    assert na == nm == 150
    assert np.array_equal(xa, xm) and np.array_equal(pa, pm) and np.array_equal(fa, fm)


def test_step_limit_stops_one_system_exactly():
    """A system with a step limit ends in the state it reaches alone after that
    many steps, however far the pool steps on; the others are unaffected."""
    from prolix_amd.engine import GBPool, Langevin

    systems = [protein(n, 160 + k, radius=synth.compact_radius(n)) for k, n in enumerate((300, 520))]
    params = [Langevin(1.0, 300.0, 1.0, seed=700 + k, safety_caps=True) for k in range(2)]
    print("[gpu] pool step limit", flush=True)
    alone = _pool_run(systems, [("admit", 0), ("steps", 70), ("read", 0)], params)
    pool = GBPool(max_atoms=2000, max_systems=4)
    slots = [pool.admit(s, p) for s, p in zip(systems, params)]
    pool.set_step_limit(slots[0], 70)
    pool.steps(50)
    pool.steps(100)  # runs past the limit of system 0
    x0, p0, f0, n0 = pool.read(slots[0])
    x1, p1, f1, n1 = pool.read(slots[1])
    pool.close()
    other = _pool_run(systems, [("admit", 1), ("steps", 150), ("read", 1)], params)
    xa, pa, fa, na = alone[0]
    assert n0 == na == 70 and n1 == 150
    assert np.array_equal(x0, xa) and np.array_equal(p0, pa) and np.array_equal(f0, fa)
    xb, pb, fb, nb = other[1]
    assert np.array_equal(x1, xb) and np.array_equal(p1, pb) and np.array_equal(f1, fb)


def test_serving_loop_small_trace():
    from prolix_amd.serving import poisson_trace, serve

    systems = [protein(int(n), 80 + k, radius=synth.compact_radius(int(n)))
               for k, n in enumerate((220, 800, 350, 1500, 260, 640))]
    reqs = poisson_trace(systems, rate_per_s=50.0, steps_range=(100, 400), seed=3)
    # a seeded 5% sample of the requests, at least two, checked against the
    # float64 oracle at the served (float32) final positions
    n_sample = max(2, int(np.ceil(0.05 * len(reqs))))
    sample = sorted(np.random.default_rng(4).choice(len(reqs), size=n_sample, replace=False).tolist())
    print("[gpu] serve", flush=True)
    results, stats = serve(reqs, chunk_steps=50, sample_energies=sample)
    assert len(results) == len(reqs)
    for r, q in zip(results, reqs):
        assert r.n_steps == q.n_steps
        assert np.isfinite(r.positions).all()
    assert stats["latency_p99_s"] >= stats["latency_p50_s"] > 0
    for k in sample:
        sysd = dict(reqs[k].system)
        x = results[k].positions.astype(np.float64)
        terms = O.energy_terms(sysd, x)
        assert results[k].energies[-1] == pytest.approx(terms["total"], rel=1e-6, abs=1e-3)
        f_ref = O.forces(sysd, x)
        keep = np.linalg.norm(f_ref, axis=1) < 9e3  # below the batched step's force cap
        err = results[k].forces[keep].astype(np.float64) - f_ref[keep]
        assert np.sqrt((err ** 2).sum() / (f_ref[keep] ** 2).sum()) <= 1e-5


def test_make_energy_fn_neighbor_argument_and_gb_cutoff():
    """A neighbour list is accepted (its contents are not needed for the value),
    and gb_cutoff=True gives the neighbour-list GB form of the golden vector."""
    from prolix_amd.physics.system import make_energy_fn
    from tests.golden import make_golden as G

    sysd, cutoff, gbc, terms, forces, born = G.load("peptide138_gbcut12")
    print("[gpu] make_energy_fn gb_cutoff", flush=True)
    fn = make_energy_fn(None, sysd, implicit_solvent=True, cutoff=cutoff, gb_cutoff=True)
    # dict systems carry no CMAP (PhysicsSystem.from_dict): compare without that term
    expected = terms[-1] - terms[G.TERMS.index("cmap")]
    nbrs = np.full((len(sysd["charges"]), 8), len(sysd["charges"]), np.int32)
    assert fn(sysd["positions"], neighbor=nbrs) == pytest.approx(expected, rel=1e-6)
    assert fn(sysd["positions"]) == fn(sysd["positions"], neighbor=nbrs)


def test_hbm_budget_admission_queues_and_keeps_results():
    """A budget that any two requests fit but not all: later arrivals wait for
    memory, no more run together than the budget holds, and every request ends
    in exactly the state it reaches with an unlimited budget (its trajectory
    does not depend on what else shares the device)."""
    from prolix_amd.serving import _bonded, hbm_bytes, poisson_trace, serve

    systems = [protein(int(n), 120 + k, radius=synth.compact_radius(int(n)))
               for k, n in enumerate((300, 700, 450, 900, 350))]
    reqs = poisson_trace(systems, rate_per_s=1000.0, steps_range=(60, 200), seed=8)
    costs = [hbm_bytes(len(r.system["charges"]), _bonded(r.system)) for r in reqs]
    budget = sorted(costs)[-1] + sorted(costs)[-2] + 1.0
    fit = max(k for k in range(1, len(costs) + 1) if sum(sorted(costs)[:k]) <= budget)
    assert fit < len(reqs)
    print("[gpu] serve with a binding HBM budget", flush=True)
    tight, st = serve(reqs, hbm_budget_bytes=budget, chunk_steps=40)
    free, _ = serve(reqs, hbm_budget_bytes=1e15, chunk_steps=40)
    # measured, not only modelled: the pool's arenas and the peak device use
    db = st["device_bytes"]
    print(f"[serve] device bytes: pool {db['pool_at_creation']}, peak {db['peak']}, budget {db['budget']:.0f}, "
          f"estimated admitted peak {db['estimated_peak_admitted']}", flush=True)
    assert 0 < db["pool_at_creation"] <= db["peak"]
    assert db["estimated_peak_admitted"] <= budget
    assert any(r.admitted - r.arrival > 1e-3 for r in tight)
    events = sorted([(r.admitted, 1) for r in tight] + [(r.finished, -1) for r in tight])
    live = peak = 0
    for _, d in events:
        live += d
        peak = max(peak, live)
    assert peak <= fit
    for a, b in zip(tight, free):
        assert a.rid == b.rid and a.n_steps == b.n_steps
        assert np.array_equal(a.positions, b.positions) and np.array_equal(a.momenta, b.momenta)
