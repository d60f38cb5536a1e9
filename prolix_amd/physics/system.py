"""Drop-in ``make_energy_fn`` backed by the gfx950 engine.

Mirrors prolix's ``physics/system.py:29-271``: same signature, same parameter
resolution, same returned callable ``total_energy(r, neighbor=None, **kw)``
(unknown keyword arguments are ignored, as in the reference) and the same
``return_decomposed`` dictionary. Energies and forces come from one engine
evaluation on the GPU; there is no host fallback.

What the reference gets from ``jax.grad(total_energy)`` is available as
``total_energy.forces(r)`` and ``total_energy.energy_and_forces(r)``
(forces = -dE/dr, analytic, from the same kernel launch).

Parameter resolution (reference line in parentheses):
- a dict system follows ``PhysicsSystem.from_dict`` (typing.py:356-425):
  radii and scaled_radii default to ones, masses to ones, CMAP fields are not
  carried over;
- an object system is read through attributes (bonds, angles, dihedrals or
  proper_dihedrals, impropers or improper_dihedrals, urey_bradley_*,
  cmap_torsions / cmap_indices / cmap_energy_grids, excl_*);
- sigmas are clamped to >= 1e-6 (system.py:107);
- ``cutoff`` in kwargs overrides ``cutoff_distance`` (system.py:134); 0 means
  no cutoff; GB is never cut;
- exclusions come from ``exclusion_spec`` (mapped to (N, 32) rows) unless the
  system carries ``excl_indices`` (system.py:161-171); without either,
  LJ/Coulomb have no exclusions (Appendix quirk Q2 of the reference);
- explicit 1-4 exceptions come from ``exclusion_spec.exception_*``;
- GB uses 78.5 / 1.0 dielectrics and the Coulomb constant 332.0637.

- ``neighbor`` (a jax_md neighbour list or its ``idx``) is accepted and its
  contents are not read: with the list, the reference restricts LJ and
  Coulomb to listed pairs closer than the cutoff (optimization.py:128-194,
  299-357) and keeps GB dense (system.py:196-216), which is the value
  computed here for every list that holds all pairs within the cutoff;
- ``gb_cutoff=True`` (keyword) selects the neighbour-list GB form of the
  batched path (generalized_born.py:266-415, batched_energy.py:999-1010):
  Born sums and GB pair terms over the pairs closer than the cutoff.

Not supported (raise): periodic boxes / PME (``use_pbc``, ``box``), the
explicit-solvent path.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from prolix_amd.physics.neighbor_list import map_exclusions_to_dense_padded

COULOMB_CONSTANT = 332.0637

_TERM_INDEX = {"bond": 0, "angle": 1, "urey_bradley": 2, "dihedral": 3, "improper": 4, "cmap": 5, "lj": 6,
               "exception": 7, "gb": 8, "coulomb": 9, "ace": 10, "total": 11}


def _np(x):
    if x is None:
        return None
    if hasattr(x, "detach"):
        x = x.detach().cpu().numpy()
    return np.asarray(x)


def _resolve_system(system, kwargs) -> dict:
    """Flatten a dict or PhysicsSystem-like object into the engine's dict."""
    if isinstance(system, dict):
        q = _np(system["charges"])
        if q is None or q.size == 0:
            raise ValueError("make_energy_fn: system['charges'] is empty — cannot infer n_atoms")
        n = q.shape[0]
        out = {
            "charges": q,
            "sigmas": _np(system["sigmas"]),
            "epsilons": _np(system["epsilons"]),
            "radii": _np(system.get("radii", np.ones(n))),
            "scaled_radii": _np(system.get("scaled_radii", np.ones(n))),
            "masses": _np(system.get("masses", np.ones(n))),
        }
        for k in ("bonds", "bond_params", "angles", "angle_params", "impropers", "improper_params",
                  "urey_bradley_bonds", "urey_bradley_params", "dihedral_params", "excl_indices",
                  "excl_scales_vdw", "excl_scales_elec"):
            if system.get(k) is not None:
                out[k] = _np(system[k])
        dih = system.get("dihedrals")
        if dih is None:
            dih = system.get("proper_dihedrals")
        if dih is not None:
            out["dihedrals"] = _np(dih)
        return out
    get = lambda name: _np(getattr(system, name, None))
    q = get("charges")
    if q is None or q.size == 0:
        raise ValueError("make_energy_fn: system.charges is empty — cannot infer n_atoms")
    n = q.shape[0]
    out = {"charges": q, "sigmas": get("sigmas"), "epsilons": get("epsilons")}
    radii = get("radii")
    out["radii"] = radii if radii is not None and radii.size else np.ones(n)
    sr = get("scaled_radii")
    out["scaled_radii"] = sr if sr is not None and sr.size else None
    m = get("masses")
    out["masses"] = m if m is not None and m.size else np.ones(n)

    def first(*names):
        for nm in names:
            v = get(nm)
            if v is not None and v.size:
                return v
        return None

    for key, names in (("bonds", ("bonds",)), ("bond_params", ("bond_params",)), ("angles", ("angles",)),
                       ("angle_params", ("angle_params",)), ("dihedrals", ("dihedrals", "proper_dihedrals")),
                       ("dihedral_params", ("dihedral_params", "proper_dihedral_params")),
                       ("impropers", ("impropers", "improper_dihedrals")),
                       ("improper_params", ("improper_params", "improper_dihedral_params")),
                       ("urey_bradley_bonds", ("urey_bradley_bonds",)),
                       ("urey_bradley_params", ("urey_bradley_params",)),
                       ("cmap_torsions", ("cmap_torsions",)), ("cmap_indices", ("cmap_indices",)),
                       ("cmap_energy_grids", ("cmap_energy_grids",)), ("excl_indices", ("excl_indices",)),
                       ("excl_scales_vdw", ("excl_scales_vdw",)), ("excl_scales_elec", ("excl_scales_elec",))):
        v = first(*names)
        if v is not None:
            out[key] = v
    if out.get("cmap_energy_grids") is not None and out["cmap_energy_grids"].ndim != 3:
        raise ValueError("cmap_energy_grids must be raw (N_maps, G, G) grids")
    return out


class _EnergyFn:
    """total_energy(r, neighbor=None, **kw) with forces from the same launch."""

    def __init__(self, engine, n_atoms):
        self._engine = engine
        self.n_atoms = n_atoms
        self._last = None

    def _eval(self, r, neighbor=None):
        key = None
        if not hasattr(r, "data_ptr"):
            a = np.asarray(r, np.float32)
            key = a.tobytes()
            if self._last is not None and self._last[0] == key:
                return self._last[1]
        e, f, b = self._engine.energy_forces(r, energies=True, born_radii=True)
        out = (e.cpu().numpy()[0], f, b)
        if key is not None:
            self._last = (key, out)
        return out

    def __call__(self, r, neighbor=None, **kwargs_run):
        return float(self._eval(r, neighbor)[0][_TERM_INDEX["total"]])

    def energy_and_forces(self, r, neighbor=None):
        e, f, _ = self._eval(r, neighbor)
        return float(e[_TERM_INDEX["total"]]), f

    def forces(self, r, neighbor=None):
        return self._eval(r, neighbor)[1]

    def terms(self, r, neighbor=None) -> dict:
        e = self._eval(r, neighbor)[0]
        return {k: float(e[i]) for k, i in _TERM_INDEX.items()}

    def born_radii(self, r, neighbor=None):
        return self._eval(r, neighbor)[2]


def make_energy_fn(displacement_fn, system, cutoff_distance: float = 9.0, return_decomposed: bool = False,
                   implicit_solvent: bool = False, exclusion_spec: Any = None, **kwargs):
    """Energy function of ``system`` on the GPU (physics/system.py:29-271)."""
    from prolix_amd.engine import GBEngine

    if kwargs.get("use_pbc", False) or kwargs.get("box") is not None:
        raise NotImplementedError("periodic boxes / PME are explicit-solvent features outside this engine")
    sysd = _resolve_system(system, kwargs)
    n = sysd["charges"].shape[0]
    if exclusion_spec is not None and sysd.get("excl_indices") is None:
        idx, sv, se = map_exclusions_to_dense_padded(exclusion_spec, max_exclusions=32)
        sysd["excl_indices"], sysd["excl_scales_vdw"], sysd["excl_scales_elec"] = idx, sv, se
    if exclusion_spec is not None and len(getattr(exclusion_spec, "exception_pairs", ())) > 0:
        sysd["exception_pairs"] = _np(exclusion_spec.exception_pairs)
        sysd["exception_sigmas"] = _np(exclusion_spec.exception_sigmas)
        sysd["exception_epsilons"] = _np(exclusion_spec.exception_epsilons)
        sysd["exception_chargeprods"] = _np(exclusion_spec.exception_chargeprods)
    cutoff = float(kwargs.get("cutoff", cutoff_distance) or 0.0)
    engine = GBEngine(sysd, cutoff=cutoff, coulomb_constant=COULOMB_CONSTANT, gb_coulomb_constant=COULOMB_CONSTANT,
                      implicit_solvent=bool(implicit_solvent), gb_cutoff=bool(kwargs.get("gb_cutoff", False)))
    fn = _EnergyFn(engine, n)
    fn.system = sysd
    if not return_decomposed:
        return fn

    def term(name):
        return lambda r, n=None: fn.terms(r, n)[name]

    def electrostatics(r, neighbor=None):
        t = fn.terms(r, neighbor)
        if implicit_solvent:
            return t["gb"], t["coulomb"], fn.born_radii(r, neighbor)
        return t["coulomb"]

    def nonpolar(r, born_radii=None):
        return fn.terms(r)["ace"] if implicit_solvent else 0.0

    return {
        "bond": term("bond"),
        "angle": term("angle"),
        "dihedral": term("dihedral"),
        "improper": term("improper"),
        "urey_bradley": term("urey_bradley"),
        "cmap": term("cmap"),
        "lj": term("lj"),
        "exception": term("exception"),
        "electrostatics": electrostatics,
        "nonpolar": nonpolar,
        "total": fn,
    }
