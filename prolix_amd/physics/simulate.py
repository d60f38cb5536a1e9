"""Drop-in integrator loop: nvt_langevin, run_minimization, run_thermalization,
run_simulation (prolix's physics/simulate.py:141-256, 357-508).

Units follow that module: A, kcal/mol, ps, masses in amu/418.4 (so that
F/m is in A/ps^2), gamma in 1/ps, kT in kcal/mol. That system is AKMA with
masses in amu and 1 ps = 1000/48.88821291839 tau, which is how the engine runs
it. The step is jax_md's nvt_langevin BAOAB (settle.py:2064-2087) on the GPU,
replayed from captured graphs; the noise is the engine's Philox4x32-10 stream
(the reference draws jax.random/threefry normals, so Langevin trajectories
agree in distribution, and exactly at gamma = 0).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

BOLTZMANN_KCAL = 0.0019872041
PS_MASS_TO_AMU = 418.4
AKMA_PS = 48.88821291839e-3  # one AKMA time unit in ps (simulate.py:56)


@dataclass
class SimulationSpec:
    """physics/simulate.py:26-52."""

    total_time_ns: float
    step_size_fs: float = 2.0
    ensemble: str = "nvt_langevin"
    temperature: float = 300.0
    pressure: float | None = None
    gamma: float = 1.0
    tau: float = 0.5
    chain_length: int = 5
    chain_steps: int = 1
    save_interval_ns: float = 0.001
    accumulate_steps: int = 500
    save_path: str = "trajectory.array_record"
    use_shake: bool = False

    @property
    def dt(self) -> float:
        return self.step_size_fs / 1000.0

    @property
    def steps(self) -> int:
        return int(self.total_time_ns * 1000.0 / self.dt)


@dataclass
class NVTLangevinState:
    """typing.py:610-628: positions (A), momentum, force (kcal/mol/A), mass, rng.

    ``momentum`` is in the units of the module: (amu/418.4) * A/ps, i.e. mass
    times velocity in A/ps. ``step`` is the number of BAOAB steps behind the
    state: with the seed in ``rng`` it fixes the position of the Philox noise
    stream, so a state holds everything the next step depends on.
    ``native`` keeps the engine's own float32 momenta (amu * A per AKMA time
    unit) next to the ``momentum`` tensor they were converted into, so that a
    state this module returned is reloaded exactly; a state whose ``momentum``
    was replaced is converted from it.
    """

    positions: object
    momentum: object
    force: object
    mass: object
    rng: int
    step: int = 0
    native: object = None

    @property
    def position(self):
        return self.positions


def _engine_of(energy_fn):
    eng = getattr(energy_fn, "_engine", None)
    if eng is None:
        raise TypeError("energy_fn must come from prolix_amd.physics.system.make_energy_fn")
    return eng


def _seed(key) -> int:
    if key is None:
        return 0
    a = np.asarray(key).reshape(-1)
    s = 0
    for v in a:
        s = (s * 0x9E3779B97F4A7C15 + int(v)) & 0xFFFFFFFFFFFFFFFF
    return s


def nvt_langevin(energy_fn, shift_fn, dt, kT, gamma=0.1, mass=1.0):
    """(init_fn, apply_fn) of the BAOAB Langevin integrator on the GPU.

    ``shift_fn`` must be free-space shifting (jax_md space.free()). apply_fn
    takes ``n_steps`` (default 1) so a loop of steps is one graph replay.
    """
    from prolix_amd.engine import Langevin

    eng = _engine_of(energy_fn)
    dt_fs = float(dt) * 1000.0
    temperature = float(kT) / BOLTZMANN_KCAL
    gamma_ps = float(gamma)

    def _masses(m):
        m = np.asarray(m, np.float64)
        if m.ndim == 0:
            m = np.full(eng.n_atoms, float(m))
        return m.reshape(-1) * PS_MASS_TO_AMU

    # The engine holds one integrator state on the device: the state object it
    # last returned and the temperature it runs at. apply_fn reloads a state it
    # did not just return (an older state, or one built by the caller), so the
    # result is a function of its ``state`` argument as in jax_md.
    held = {"state": None, "kT": None}

    def _load(state, kT_now):
        if state.native is not None and state.native[1] is state.momentum:
            p = state.native[0]
        else:  # (amu/418.4) A/ps -> amu A/tau: a factor 1/AKMA_PS (418.4 AKMA_PS^2 = 1)
            p = eng.torch.as_tensor(state.momentum, dtype=eng.torch.float64, device=eng.device) / AKMA_PS
        eng.init(state.positions, Langevin(dt_fs, float(kT_now) / BOLTZMANN_KCAL, gamma_ps, seed=int(state.rng)),
                 momenta=p)
        eng.set_step_counts(int(state.step))

    def _wrap(mass_, seed, step, kT_now):
        x, p, f = eng.state()
        mom = (p.double() * AKMA_PS).float()
        st = NVTLangevinState(x, mom, f, mass_, seed, step, native=(p, mom))
        held["state"], held["kT"] = st, float(kT_now)
        return st

    def init_fn(key, R, mass=mass, **kwargs):
        _kT = kwargs.get("kT", kT)
        eng.set_masses(_masses(mass))
        seed = _seed(key)
        eng.init(R, Langevin(dt_fs, float(_kT) / BOLTZMANN_KCAL, gamma_ps, seed=seed))
        return _wrap(mass, seed, 0, _kT)

    def apply_fn(state, n_steps: int = 1, **kwargs):
        _kT = float(kwargs.get("kT", kT))
        if held["state"] is not state or held["kT"] != _kT:
            _load(state, _kT)
        eng.steps(int(n_steps))
        return _wrap(state.mass, state.rng, int(state.step) + int(n_steps), _kT)

    return init_fn, apply_fn


def run_minimization(energy_fn, initial_positions, steps: int = 500, dt_start: float = 2e-3, dt_max: float = 4e-3,
                     max_displacement_per_step: float = 0.1, sd_steps: int = 200):
    """Steepest descent with 4-point backtracking (physics/simulate.py:141-196)."""
    from prolix_amd.minimize import minimize

    eng = _engine_of(energy_fn)
    return minimize(eng, initial_positions, steps=sd_steps, dt_max=dt_max,
                    max_displacement=max_displacement_per_step)


def run_thermalization(energy_fn, initial_positions, temperature: float = 300.0, steps: int = 1000,
                       dt: float = 2e-3, gamma: float = 1.0, mass=1.0, use_shake: bool = False,
                       constraints=None, key=None):
    """NVT Langevin for ``steps`` steps, returns final positions (physics/simulate.py:199-256)."""
    if use_shake and constraints is not None:
        raise NotImplementedError("RATTLE/SHAKE constraints are not part of this engine")
    init_fn, apply_fn = nvt_langevin(energy_fn, None, dt=dt, kT=BOLTZMANN_KCAL * temperature, gamma=gamma,
                                     mass=mass)
    state = init_fn(key, initial_positions)
    if steps > 0:
        state = apply_fn(state, n_steps=steps)
        from prolix_amd.checks import check_engine_health

        check_engine_health(_engine_of(energy_fn), stage="thermalization")
    return state.position


def run_simulation(system_params, initial_positions, temperature: float = 300.0, min_steps: int = 500,
                   therm_steps: int = 1000, dielectric_constant: float = 1.0, implicit_solvent: bool = True,
                   solvent_dielectric: float = 78.5, solute_dielectric: float = 1.0, key=None, sim_spec=None,
                   exclusion_spec=None):
    """Minimise -> thermalise -> optional Langevin production (physics/simulate.py:357-508).

    As in the reference, the dielectric keyword arguments do not reach the GB
    term (it uses 78.5 / 1.0), masses are taken as amu/418.4 and the
    thermalisation runs 2 fs steps at gamma = 1/ps. The reference's
    ``Protein`` NameError (physics/simulate.py:16-19, 402) is not reproduced:
    any system accepted by make_energy_fn works.
    """
    from prolix_amd.physics.system import make_energy_fn

    temperature_prod = temperature
    dt_production = 2e-3
    steps_production = 0
    if sim_spec is not None:
        temperature_prod = sim_spec.temperature
        dt_production = sim_spec.dt
        steps_production = sim_spec.steps
    energy_fn = make_energy_fn(None, system_params, implicit_solvent=implicit_solvent,
                               exclusion_spec=exclusion_spec)
    from prolix_amd.checks import check_initial_energy

    check_initial_energy(_engine_of(energy_fn), initial_positions)  # simulate.py:571-582
    positions = run_minimization(energy_fn, initial_positions, steps=min_steps)
    masses_raw = system_params.get("masses") if isinstance(system_params, dict) else getattr(
        system_params, "masses", None)
    masses = np.asarray(masses_raw, np.float64) / PS_MASS_TO_AMU if masses_raw is not None else 1.0
    r_therm = run_thermalization(energy_fn, positions, temperature=temperature_prod, steps=therm_steps,
                                 mass=masses, key=key)
    if sim_spec is None or steps_production <= 0:
        return r_therm
    if sim_spec.ensemble != "nvt_langevin":
        raise NotImplementedError(f"ensemble {sim_spec.ensemble!r}: only nvt_langevin runs on this engine")
    return run_thermalization(energy_fn, r_therm, temperature=temperature_prod, steps=steps_production,
                              dt=dt_production, gamma=sim_spec.gamma, mass=masses, key=key)
