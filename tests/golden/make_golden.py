"""Write the golden vectors under tests/golden/ from the float64 oracle.

    python -m tests.golden.make_golden

Each case is one ``<name>.npz`` (plain arrays, no pickles): the inputs under
``in_<key>`` (every float exactly representable in float32, so the float32
device path and the float64 oracle see the same numbers), and the expected
outputs ``out_terms`` (the 12 energy terms of the engine's term order,
kcal/mol), ``out_forces`` (N, 3) kcal/mol/A and ``out_born`` (N,) A, plus the
evaluation options ``opt_cutoff`` and ``opt_gb_cutoff``.

Cases:
- peptide138: the 10-residue, 138-atom synthetic peptide (seed 0) with
  1-2/1-3 exclusions, 1-4 scaling and CMAP on its 8 phi/psi pairs, no cutoff;
- peptide138_cut9: the same system with the drop-in default 9 A LJ/Coulomb cutoff;
- chain8: a hand-built 8-atom chain carrying one of every bonded kind
  (bond, Urey-Bradley, angle, two-term proper torsion, periodic and
  harmonic improper, explicit 1-4 exception) next to GB/Coulomb/LJ;
- ions3: three free ions (GB, Coulomb and LJ only);
- peptide138_gbcut12, chain8_gbcut4: the neighbour-list GB form (Born sums
  and GB pairs over the pairs closer than the cutoff) at 12 A and 4 A.

No pair of any cutoff case lies within 5e-5 A of its cutoff (100 times the
float32 rounding of r near 10 A), so float32 and
float64 evaluations classify every pair the same way.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

from oracle import gb_oracle as O
from prolix_amd import synth
from prolix_amd.physics.neighbor_list import ExclusionSpec, map_exclusions_to_dense_padded

HERE = Path(__file__).resolve().parent

TERMS = ("bond", "angle", "urey_bradley", "dihedral", "improper", "cmap", "lj", "exception", "gb", "coulomb",
         "ace", "total")


def _f32(a):
    return np.asarray(np.asarray(a, np.float64).astype(np.float32), np.float64)


def peptide() -> dict:
    p = synth.peptide_138(0)
    d = p.as_dict()
    spec = ExclusionSpec.from_bonds(d["bonds"], p.n_atoms)
    d["excl_indices"], d["excl_scales_vdw"], d["excl_scales_elec"] = map_exclusions_to_dense_padded(spec)
    return d


def chain8() -> dict:
    x = np.array([[0.0, 0.0, 0.0], [1.5, 0.1, 0.0], [2.1, 1.4, 0.2], [3.6, 1.5, -0.3],
                  [4.2, 2.8, 0.4], [5.7, 2.6, 0.9], [2.3, -0.9, 1.1], [6.3, 3.9, 1.6]])
    return {
        "positions": x,
        "charges": np.array([0.3, -0.25, 0.1, -0.4, 0.35, -0.2, 0.15, -0.05]),
        "sigmas": np.array([1.9, 3.2, 2.4, 1.1, 2.9, 3.4, 1.5, 2.2]),
        "epsilons": np.array([0.05, 0.15, 0.1, 0.02, 0.2, 0.08, 0.12, 0.03]),
        "radii": np.array([1.5, 1.7, 1.3, 1.9, 1.55, 1.8, 1.2, 1.45]),
        "scaled_radii": np.array([1.2, 1.3, 1.0, 1.5, 1.25, 1.4, 0.95, 1.15]),
        "masses": np.array([14.0, 12.0, 12.0, 16.0, 14.0, 12.0, 1.0, 16.0]),
        "bonds": np.array([[0, 1], [1, 2], [2, 3], [3, 4], [4, 5], [1, 6], [5, 7]]),
        "bond_params": np.array([[1.45, 300.0], [1.5, 250.0], [1.55, 310.0], [1.4, 280.0], [1.5, 260.0],
                                 [1.1, 340.0], [1.3, 400.0]]),
        "urey_bradley_bonds": np.array([[0, 2]]),
        "urey_bradley_params": np.array([[2.4, 20.0]]),
        "angles": np.array([[0, 1, 2], [1, 2, 3], [2, 3, 4], [3, 4, 5], [0, 1, 6], [4, 5, 7]]),
        "angle_params": np.array([[1.9, 60.0], [2.0, 55.0], [1.95, 70.0], [2.05, 50.0], [1.85, 35.0],
                                  [2.1, 45.0]]),
        "dihedrals": np.array([[0, 1, 2, 3], [1, 2, 3, 4], [2, 3, 4, 5]]),
        "dihedral_params": np.array([[[1.0, 0.0, 0.8], [3.0, 0.5, 0.2]], [[2.0, 3.14159, 1.1], [1.0, 0.3, 0.4]],
                                     [[3.0, 0.0, 0.15], [2.0, 1.2, 0.6]]]),
        "impropers": np.array([[6, 0, 1, 2]]),
        "improper_params": np.array([[[2.0, 3.14159, 1.1]]]),
        "exception_pairs": np.array([[0, 3], [6, 3]]),
        "exception_sigmas": np.array([2.1, 1.8]),
        "exception_epsilons": np.array([0.06, 0.04]),
        "exception_chargeprods": np.array([-0.06, -0.03]),
        "excl_indices": np.array([[1, 2, -1, -1], [0, 2, 6, -1], [0, 1, 3, 6], [2, 4, -1, -1],
                                  [3, 5, -1, -1], [4, 7, -1, -1], [1, 2, -1, -1], [5, -1, -1, -1]]),
        "excl_scales_vdw": np.zeros((8, 4)),
        "excl_scales_elec": np.zeros((8, 4)),
    }


def chain8_harmonic() -> dict:
    d = chain8()
    d["improper_params"] = np.array([[[25.0, 0.35]]])  # harmonic [k, phi0]
    # one row-asymmetric entry: atom 0 lists 4 with 1-4 scales, atom 4 does not
    # list 0, so the pair carries the mean of the two rows' scales
    d["excl_indices"][0, 2] = 4
    d["excl_scales_vdw"][0, 2] = 0.5
    d["excl_scales_elec"][0, 2] = 0.8333333333
    return d


def ions3() -> dict:
    return {
        "positions": np.array([[0.0, 0.0, 0.0], [4.1, 0.0, 0.0], [1.3, 3.7, -0.8]]),
        "charges": np.array([1.0, -1.0, 0.5]),
        "sigmas": np.array([2.0, 3.0, 2.5]),
        "epsilons": np.array([0.1, 0.15, 0.05]),
        "radii": np.array([1.8, 2.0, 1.5]),
        "scaled_radii": np.array([1.44, 1.6, 1.2]),
        "masses": np.array([23.0, 35.5, 40.0]),
    }


FLOAT_KEYS = ("positions", "charges", "sigmas", "epsilons", "radii", "scaled_radii", "masses", "bond_params",
              "urey_bradley_params", "angle_params", "dihedral_params", "improper_params", "cmap_energy_grids",
              "excl_scales_vdw", "excl_scales_elec", "exception_sigmas", "exception_epsilons",
              "exception_chargeprods")


def evaluate(sysd: dict, cutoff: float = 0.0, gb_cutoff: bool = False):
    kw = {"cutoff": cutoff}
    if gb_cutoff:
        kw["gb_cutoff"] = True
    t = O.energy_terms(sysd, sysd["positions"], **kw)
    f = O.forces(sysd, sysd["positions"], **kw)
    return np.array([t[k] for k in TERMS]), f, np.asarray(t["_born_radii"])


def cases():
    return {
        "peptide138": (peptide(), 0.0, False),
        "peptide138_cut9": (peptide(), 9.0, False),
        "chain8": (chain8(), 0.0, False),
        "chain8_harmonic": (chain8_harmonic(), 0.0, False),
        "ions3": (ions3(), 0.0, False),
        "peptide138_gbcut12": (peptide(), 12.0, True),
        "chain8_gbcut4": (chain8(), 4.0, True),
    }


def load(name: str):
    """(system dict, cutoff, gb_cutoff, terms, forces, born) of a stored case."""
    z = np.load(HERE / f"{name}.npz")
    sysd = {k[3:]: z[k] for k in z.files if k.startswith("in_")}
    return (sysd, float(z["opt_cutoff"]), bool(z["opt_gb_cutoff"]), z["out_terms"], z["out_forces"],
            z["out_born"])


def main():
    for name, (sysd, cutoff, gbc) in cases().items():
        sysd = {k: v for k, v in sysd.items() if v is not None and k != "elements"}
        for k in FLOAT_KEYS:
            if k in sysd:
                sysd[k] = _f32(sysd[k])
        if cutoff > 0:
            x = sysd["positions"]
            r = np.linalg.norm(x[:, None] - x[None], axis=-1)
            assert not np.any(np.abs(r - cutoff) < 5e-5), f"{name}: a pair sits on the cutoff"
        terms, f, born = evaluate(sysd, cutoff, gbc)
        arrays = {f"in_{k}": np.asarray(v) for k, v in sysd.items()}
        np.savez_compressed(HERE / f"{name}.npz", **arrays, out_terms=terms, out_forces=f, out_born=born,
                            opt_cutoff=np.float64(cutoff), opt_gb_cutoff=np.int32(gbc))
        print(f"{name}: n={len(sysd['charges'])} total={terms[-1]:.10g}")


if __name__ == "__main__":
    main()
