"""Python handle on the native engine: packed systems, energies/forces, BAOAB.

Device memory and streams come from PyTorch (ROCm build): every per-step
array is a ``torch.float32`` CUDA tensor whose pointer is handed to the C ABI,
and work is ordered on the current torch stream. Nothing here computes
physics; it only moves buffers and calls the HIP kernels.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from prolix_amd import _native as N


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("prolix_amd's engine needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


def _stream(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _vp(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


@dataclass
class Langevin:
    """BAOAB parameters of one system (fs, K, 1/ps, Philox seed)."""

    dt_fs: float = 1.0
    temperature: float = 300.0
    gamma_ps: float = 1.0
    seed: int = 0
    safety_caps: bool = False  # the batched step's velocity / displacement / force caps

    def c(self) -> N.LangevinParams:
        return N.LangevinParams(float(self.dt_fs), float(self.temperature), float(self.gamma_ps),
                                int(self.seed) & 0xFFFFFFFFFFFFFFFF, 1 if self.safety_caps else 0, 0)


class GBEngine:
    """Many systems packed into one engine (a single system is a pack of one)."""

    def __init__(self, systems, *, cutoff: float = 0.0, coulomb_constant: float = 332.0637,
                 gb_coulomb_constant: float = 332.0637, solute_dielectric: float = 1.0,
                 solvent_dielectric: float = 78.5, dielectric_offset: float = 0.09,
                 implicit_solvent: bool = True, gb_cutoff: bool = False):
        """``cutoff`` (A, 0 = none) applies to LJ and Coulomb; with ``gb_cutoff``
        the Born sums and GB pair terms also run over the pairs closer than the
        cutoff only (the neighbour-list GB form)."""
        torch = _torch()
        self.torch = torch
        if isinstance(systems, dict):
            systems = [systems]
        opts = dict(cutoff=cutoff, coulomb_constant=coulomb_constant, gb_coulomb_constant=gb_coulomb_constant,
                    solute_dielectric=solute_dielectric, solvent_dielectric=solvent_dielectric,
                    dielectric_offset=dielectric_offset, implicit_solvent=implicit_solvent, gb_cutoff=gb_cutoff)
        self._descs = [N.Desc(s, **opts) for s in systems]
        arr = (N.SystemDesc * len(self._descs))(*[d.desc for d in self._descs])
        h = ctypes.c_void_p()
        N.check(N.lib().prx_system_create(arr, len(self._descs), ctypes.byref(h)), "prx_system_create")
        self._h = h
        self.n_systems = len(self._descs)
        off = (ctypes.c_int32 * (self.n_systems + 1))()
        N.check(N.lib().prx_engine_layout(h, off, self.n_systems + 1), "prx_engine_layout")
        self.offsets = np.array(off[:], dtype=np.int64)
        self.n_atoms = int(self.offsets[-1])
        self.sizes = np.diff(self.offsets)
        self.device = torch.device("cuda")

    # ---- helpers --------------------------------------------------------
    def pack(self, positions_list):
        """Concatenate per-system (n_s, 3) arrays into one device tensor."""
        t = self.torch
        parts = [t.as_tensor(np.asarray(p, np.float32).reshape(-1, 3)) for p in positions_list]
        return t.cat(parts).to(self.device).contiguous()

    def unpack(self, packed):
        a = packed.detach().cpu().numpy() if hasattr(packed, "detach") else np.asarray(packed)
        return [a[self.offsets[s]:self.offsets[s + 1]] for s in range(self.n_systems)]

    def _pos_tensor(self, positions):
        t = self.torch
        if isinstance(positions, (list, tuple)):
            return self.pack(positions)
        x = t.as_tensor(positions, dtype=t.float32, device=self.device).reshape(-1, 3).contiguous()
        if x.shape[0] != self.n_atoms:
            raise ValueError(f"positions have {x.shape[0]} atoms, engine has {self.n_atoms}")
        return x

    # ---- energy / forces ------------------------------------------------
    def energy_forces(self, positions, *, energies: bool = True, born_radii: bool = False):
        """Returns (energies (n_systems, 12) float64 | None, forces (N, 3) float32, born | None)."""
        t = self.torch
        x = self._pos_tensor(positions)
        f = t.empty((self.n_atoms, 3), dtype=t.float32, device=self.device)
        e = t.empty((self.n_systems, N.N_ENERGY), dtype=t.float64, device=self.device) if energies else None
        b = t.empty((self.n_atoms,), dtype=t.float32, device=self.device) if born_radii else None
        N.check(N.lib().prx_energy_forces(self._h, _vp(x), _vp(f), _vp(e), _vp(b), _stream(t)), "prx_energy_forces")
        return e, f, b

    # ---- integrator -----------------------------------------------------
    def init(self, positions, params, momenta=None):
        t = self.torch
        if isinstance(params, Langevin):
            params = [params] * self.n_systems
        if len(params) != self.n_systems:
            raise ValueError("one Langevin parameter set per system is required")
        x = self._pos_tensor(positions)
        m = None if momenta is None else self._pos_tensor(momenta)
        arr = (N.LangevinParams * self.n_systems)(*[p.c() for p in params])
        N.check(N.lib().prx_integrator_init(self._h, _vp(x), _vp(m), arr, _stream(t)), "prx_integrator_init")

    def set_masses(self, masses_amu):
        """Per-atom masses (packed, amu) used by the integrator from the next init."""
        m = np.ascontiguousarray(np.asarray(masses_amu, np.float64).reshape(-1))
        if m.shape[0] != self.n_atoms:
            raise ValueError("one mass per atom is required")
        N.check(N.lib().prx_engine_set_masses(self._h, m.ctypes.data_as(N.c_f64p)), "prx_engine_set_masses")

    def steps(self, n_steps: int):
        N.check(N.lib().prx_integrator_steps(self._h, int(n_steps), _stream(self.torch)), "prx_integrator_steps")

    def set_step_counts(self, steps):
        """Set each system's step count, which is the position of its Philox noise
        stream (after ``init`` with saved momenta: resume a saved trajectory)."""
        if np.ndim(steps) == 0:
            steps = [int(steps)] * self.n_systems
        if len(steps) != self.n_systems:
            raise ValueError("one step count per system is required")
        arr = (ctypes.c_int64 * self.n_systems)(*[int(x) for x in steps])
        N.check(N.lib().prx_integrator_set_step_counts(self._h, arr, _stream(self.torch)),
                "prx_integrator_set_step_counts")

    def nonfinite_counts(self) -> np.ndarray:
        """Per system: atom-steps since init that left a non-finite position or
        momentum (0 for a healthy system)."""
        arr = (ctypes.c_int64 * max(1, self.n_systems))()
        N.check(N.lib().prx_integrator_health(self._h, arr, _stream(self.torch)), "prx_integrator_health")
        return np.array(arr[: self.n_systems], np.int64)

    def set_nl_skin(self, skin: float):
        """Verlet skin (A) of the neighbour lists of gb_cut systems: 0 rebuilds
        the list at every evaluation, skin > 0 when an atom moved > skin / 2."""
        N.check(N.lib().prx_engine_set_nl_skin(self._h, float(skin)), "prx_engine_set_nl_skin")

    STEP_MODES = {"auto": 0, "graph": 1, "persistent": 2, "fused": 3}

    def set_step_mode(self, mode: str):
        """How steps run, with the same results bit for bit: "graph" (captured
        per-pass kernels), "persistent" (one cooperative kernel per 1,000
        steps, grid-wide barriers), "fused" (one workgroup per system,
        workgroup barriers only; not for neighbour-list systems) or "auto"
        (the default: fused when every system has at most 192 atoms and none
        uses a neighbour list, graphs otherwise; the fastest on MI355X)."""
        N.check(N.lib().prx_engine_set_step_mode(self._h, self.STEP_MODES[mode]), "prx_engine_set_step_mode")

    def set_fork_min_atoms(self, min_atoms: int):
        """Run the bonded kernel beside the pair passes from this many atoms on."""
        N.check(N.lib().prx_engine_set_fork_min_atoms(self._h, int(min_atoms)), "prx_engine_set_fork_min_atoms")

    def state(self):
        t = self.torch
        x = t.empty((self.n_atoms, 3), dtype=t.float32, device=self.device)
        p = t.empty_like(x)
        f = t.empty_like(x)
        N.check(N.lib().prx_integrator_state(self._h, _vp(x), _vp(p), _vp(f), _stream(t)), "prx_integrator_state")
        return x, p, f

    def kinetic(self):
        t = self.torch
        ke = t.empty((self.n_systems,), dtype=t.float64, device=self.device)
        N.check(N.lib().prx_integrator_kinetic(self._h, _vp(ke), _stream(t)), "prx_integrator_kinetic")
        return ke

    def recenter_com(self):
        """Subtract each system's mass-weighted centre of mass from its positions
        (batched_simulate.py:58-81); momenta and forces are unchanged."""
        N.check(N.lib().prx_recenter_com(self._h, _stream(self.torch)), "prx_recenter_com")

    def step_count(self, system: int = 0) -> int:
        c = ctypes.c_int64()
        N.check(N.lib().prx_integrator_step_count(self._h, int(system), ctypes.byref(c)), "prx_integrator_step_count")
        return c.value

    PASSES = ("born_pairs", "born_gather", "force_pairs", "force_gather", "chain_pairs", "bonded", "chain_gather")

    def profile_passes(self, n_reps: int = 10) -> dict:
        """Mean HIP-event duration (ms) of each force pass launched on its own."""
        ms = (ctypes.c_double * 7)()
        N.check(N.lib().prx_profile_passes(self._h, int(n_reps), ms), "prx_profile_passes")
        return dict(zip(self.PASSES, list(ms)))

    def profile_steps(self, n_steps: int = 20) -> dict:
        """Mean HIP-event duration (ms) of each force pass inside the step graph,
        over n_steps real integrator steps (the integrator advances)."""
        ms = (ctypes.c_double * 7)()
        N.check(N.lib().prx_profile_steps(self._h, int(n_steps), ms), "prx_profile_steps")
        return dict(zip(self.PASSES, list(ms)))

    def work_items(self) -> int:
        n = ctypes.c_int64()
        N.check(N.lib().prx_engine_work_items(self._h, ctypes.byref(n)), "prx_engine_work_items")
        return n.value

    def warnings(self, system: int = 0):
        """Safety-cap event counts (velocity, force, constraint, displacement)."""
        c = (ctypes.c_int64 * 4)()
        N.check(N.lib().prx_integrator_warnings(self._h, int(system), c), "prx_integrator_warnings")
        return list(c)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.torch.cuda.synchronize()
            N.check(N.lib().prx_engine_destroy(self._h), "prx_engine_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GBPool:
    """A serving pool: systems are admitted and released while others keep stepping."""

    def __init__(self, max_atoms: int, max_systems: int = 1024, max_bonded: int | None = None,
                 max_cmap_coeff: int | None = None, max_excl_tiles: int | None = None):
        torch = _torch()
        self.torch = torch
        lim = N.PoolLimits(int(max_atoms), int(max_bonded or 12 * max_atoms),
                           int(max_cmap_coeff or max(1, max_systems) * 4 * 24 * 24 * 16),
                           int(max_excl_tiles or max(64, max_atoms // 16)), int(max_systems))
        h = ctypes.c_void_p()
        N.check(N.lib().prx_pool_create(ctypes.byref(lim), ctypes.byref(h)), "prx_pool_create")
        self._h = h
        self._descs = {}
        self.sizes = {}

    @staticmethod
    def describe(sysd: dict, **opts):
        """The host-side description of a system that admit takes (no GPU
        work): made ahead, for example while the pool steps."""
        return N.Desc(sysd, **opts)

    def admit(self, sysd: dict, params: Langevin, positions=None, desc=None, **opts) -> int:
        """Upload one system; returns its slot. Its momenta and forces are
        computed at the next step call or read, together with every other
        system admitted since the last one. desc: the system's describe()
        result, if made ahead."""
        d = desc if desc is not None else N.Desc(sysd, **opts)
        x = np.ascontiguousarray(np.asarray(sysd["positions"] if positions is None else positions, np.float32))
        slot = ctypes.c_int32()
        p = params.c()
        code = N.lib().prx_pool_admit(self._h, ctypes.byref(d.desc), x.ctypes.data_as(N.c_f32p), ctypes.byref(p),
                                      ctypes.byref(slot), _stream(self.torch))
        if code == 3:  # PRX_ERR_CAPACITY
            return -1
        N.check(code, "prx_pool_admit")
        self._descs[slot.value] = d
        self.sizes[slot.value] = d.n_atoms
        return slot.value

    def set_step_mode(self, mode: str):
        """As GBEngine.set_step_mode, for the pool's live systems."""
        N.check(N.lib().prx_engine_set_step_mode(self._h, GBEngine.STEP_MODES[mode]), "prx_engine_set_step_mode")

    def set_step_limit(self, slot: int, limit: int):
        """Stop the system after `limit` steps in all (0: never); later step
        calls leave its state as it was after that step."""
        N.check(N.lib().prx_pool_set_step_limit(self._h, int(slot), int(limit), _stream(self.torch)),
                "prx_pool_set_step_limit")

    def release(self, slot: int):
        N.check(N.lib().prx_pool_release(self._h, int(slot), _stream(self.torch)), "prx_pool_release")
        self._descs.pop(slot, None)
        self.sizes.pop(slot, None)

    def steps(self, n_steps: int):
        N.check(N.lib().prx_integrator_steps(self._h, int(n_steps), _stream(self.torch)), "prx_integrator_steps")

    def read(self, slot: int):
        n = self.sizes[slot]
        x = np.zeros((n, 3), np.float32)
        p = np.zeros((n, 3), np.float32)
        f = np.zeros((n, 3), np.float32)
        c = ctypes.c_int64()
        N.check(N.lib().prx_pool_read(self._h, int(slot), x.ctypes.data_as(N.c_f32p), p.ctypes.data_as(N.c_f32p),
                                      f.ctypes.data_as(N.c_f32p), ctypes.byref(c), _stream(self.torch)),
                "prx_pool_read")
        return x, p, f, c.value

    def energies(self, slot: int):
        e = np.zeros(N.N_ENERGY)
        N.check(N.lib().prx_pool_energy(self._h, int(slot), e.ctypes.data_as(N.c_f64p), _stream(self.torch)),
                "prx_pool_energy")
        return e

    def nonfinite_counts(self) -> dict:
        """Per live slot: atom-steps since admission that left a non-finite
        position or momentum."""
        live = self.live()
        arr = (ctypes.c_int64 * max(1, len(live)))()
        N.check(N.lib().prx_integrator_health(self._h, arr, _stream(self.torch)), "prx_integrator_health")
        return {slot: int(arr[k]) for k, slot in enumerate(live)}

    def live(self):
        n = ctypes.c_int32()
        buf = (ctypes.c_int32 * 65536)()
        N.check(N.lib().prx_pool_slots(self._h, buf, 65536, ctypes.byref(n)), "prx_pool_slots")
        return list(buf[: n.value])

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.torch.cuda.synchronize()
            N.check(N.lib().prx_engine_destroy(self._h), "prx_engine_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
