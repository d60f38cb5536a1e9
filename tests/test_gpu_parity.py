"""HIP path vs the float64 oracle: Born radii, per-term energies, forces.

Tolerances: energy relative error <= 1e-6 and force RMS relative error
<= 1e-5 against the float64 oracle on the same float32-rounded inputs. The
energy error is taken relative to max(|E_total|, 0.2 * sum_k |E_k|), which is
|E_total| for every benchmark configuration (peptide: 1574 of 2731; globule:
27561 of 36302) and only differs for small systems whose total is a near
cancellation of its terms.
"""

from __future__ import annotations

import json
import os
import sys

import numpy as np
import pytest

from oracle import gb_oracle as O
from prolix_amd import _native as N
from prolix_amd import synth
from tests.helpers import f32_round, protein, rel_rms, with_exclusions

pytestmark = pytest.mark.gpu

E_TOL = 1e-6
F_TOL = 1e-5
_SUMMARY = {}


def _note(case, **kw):
    _SUMMARY[case] = kw
    out = os.environ.get("PRX_PARITY_OUT")
    if out:
        os.makedirs(os.path.dirname(out), exist_ok=True)
        with open(out, "w") as fh:
            json.dump(_SUMMARY, fh, indent=1, default=float)


def _engine(sysd, **kw):
    from prolix_amd.engine import GBEngine

    return GBEngine(sysd, **kw)


def _gpu_eval(sysd, **kw):
    print(f"[gpu] energy_forces n={len(sysd['charges'])} {kw}", flush=True)
    eng = _engine(sysd, **kw)
    e, f, b = eng.energy_forces(sysd["positions"], energies=True, born_radii=True)
    out = (e.cpu().numpy()[0], f.cpu().numpy(), b.cpu().numpy())
    eng.close()
    return out


def _oracle(sysd, **kw):
    terms = O.energy_terms(sysd, sysd["positions"], **kw)
    forces = O.forces(sysd, sysd["positions"], **kw)
    return terms, forces


def _compare(case, sysd, okw=None, gkw=None):
    okw = okw or {}
    gkw = gkw or {}
    e, f, b = _gpu_eval(sysd, **gkw)
    terms, fo = _oracle(sysd, **okw)
    scale = sum(abs(terms[k]) for k in O.TERM_NAMES)
    errs = {}
    for k, name in enumerate(N.TERMS[:-1]):
        errs[name] = abs(e[k] - terms[name])
    # Relative to the energy scale: the total, or a fifth of the summed term
    # magnitudes where the total is a cancellation of large opposite terms
    # (float32 pair arithmetic resolves each term, not their difference).
    e_scale = max(abs(terms["total"]), 0.2 * scale, 1e-30)
    tot_rel = abs(e[-1] - terms["total"]) / e_scale
    f_rel = rel_rms(f, fo)
    info = dict(n=len(sysd["charges"]), total_gpu=float(e[-1]), total_oracle=terms["total"], total_rel=tot_rel,
                force_rel_rms=f_rel, term_abs_err=errs, term_scale=scale)
    if "_born_radii" in terms:
        info["born_rel_max"] = float(np.max(np.abs(b - terms["_born_radii"]) / terms["_born_radii"]))
    _note(case, **info)
    print(f"[{case}] total rel {tot_rel:.2e}  force rel-rms {f_rel:.2e}  born {info.get('born_rel_max', 0):.2e}",
          flush=True)
    return e, f, b, terms, fo, info


def test_born_ion_closed_form():
    """Single ion: B = R - 0.09 and E_GB = -1/2 k (1/eps_in - 1/eps_out) q^2 / B
    (tests/physics/test_generalized_born.py:11-52)."""
    sysd = {"positions": np.zeros((1, 3)), "charges": np.array([1.0]), "sigmas": np.array([1.0]),
            "epsilons": np.array([0.0]), "radii": np.array([2.0])}
    e, f, b = _gpu_eval(sysd)
    B = 2.0 - 0.09
    expected = -0.5 * 332.0637 * (1.0 - 1.0 / 78.5) / B
    assert abs(b[0] - B) / B < 1e-6
    assert abs(e[N.TERMS.index("gb")] - expected) / abs(expected) < 1e-6
    assert np.all(f == 0.0)


def test_two_atoms_finite_difference_geometry():
    """The reference's 2-particle GB gradient case (test_generalized_born.py:54-91)."""
    sysd = f32_round({"positions": np.array([[0.0, 0.0, 0.0], [3.5, 0.0, 0.0]]), "charges": np.array([1.0, -1.0]),
                      "sigmas": np.array([1.0, 1.0]), "epsilons": np.array([0.0, 0.0]),
                      "radii": np.array([1.5, 1.5])})
    _, f, _, terms, fo, info = _compare("two_atoms", sysd)
    assert info["total_rel"] < E_TOL
    # This geometry's net force (-0.28) is 1% of its parts (Coulomb +27.1, GB
    # direct -23.9, Born chain -3.5): float32 pair arithmetic resolves each
    # part to ~1e-7, so the error is bounded against the component scale.
    f_coul = O.forces(sysd, sysd["positions"], implicit_solvent=False)
    comp = np.sqrt(np.mean(f_coul**2) + np.mean((fo - f_coul) ** 2))
    err = np.sqrt(np.mean((f - fo) ** 2))
    print(f"[two_atoms] force error / component scale {err / comp:.2e}", flush=True)
    assert err / comp < F_TOL
    assert abs(f[0, 0] + f[1, 0]) <= 1e-6 * comp


@pytest.mark.parametrize("n_atoms,seed", [(40, 3), (64, 4), (65, 5), (138, 0), (300, 6)])
def test_small_proteins(n_atoms, seed):
    sysd = protein(n_atoms, seed)
    _, _, _, _, _, info = _compare(f"protein_{n_atoms}", sysd)
    if n_atoms == 138:  # the 138-atom peptide: the error is relative to |E_total| itself
        assert abs(info["total_oracle"]) >= 0.2 * info["term_scale"]
    assert info["total_rel"] < E_TOL
    assert info["force_rel_rms"] < F_TOL
    assert info["born_rel_max"] < 1e-5
    for name, err in info["term_abs_err"].items():
        assert err <= E_TOL * info["term_scale"] + 1e-9, (name, err)


def test_globule_2500():
    sysd = protein(2500, 1, radius=synth.compact_radius(2500))
    _, _, _, _, _, info = _compare("globule_2500", sysd)
    assert abs(info["total_oracle"]) >= 0.2 * info["term_scale"]  # relative to |E_total| itself
    assert info["total_rel"] < E_TOL
    assert info["force_rel_rms"] < F_TOL
    assert info["born_rel_max"] < 1e-5


def test_cutoff_and_no_exclusions():
    """Default make_energy_fn cutoff (9 A) on LJ/Coulomb; GB stays uncut."""
    sysd = protein(300, 11)
    _, _, _, _, _, info = _compare("cutoff9", sysd, okw={"cutoff": 9.0}, gkw={"cutoff": 9.0})
    assert info["total_rel"] < E_TOL
    assert info["force_rel_rms"] < F_TOL
    bare = protein(200, 12, exclusions=False)
    _, _, _, _, _, info2 = _compare("no_exclusions", bare)
    assert info2["force_rel_rms"] < F_TOL


def test_harmonic_impropers_exceptions_ub():
    sysd = protein(150, 13)
    rng = np.random.default_rng(0)
    imp = sysd["impropers"]
    sysd["improper_params"] = np.stack([rng.uniform(1, 5, len(imp)), rng.uniform(-0.3, 0.3, len(imp))], 1)[:, None, :]
    sysd["urey_bradley_bonds"] = sysd["angles"][:20, [0, 2]]
    d = np.linalg.norm(sysd["positions"][sysd["urey_bradley_bonds"][:, 0]]
                       - sysd["positions"][sysd["urey_bradley_bonds"][:, 1]], axis=1)
    sysd["urey_bradley_params"] = np.stack([d + 0.05, np.full(20, 30.0)], 1)
    pairs = np.array([[0, 40], [3, 77], [10, 120]], np.int32)
    sysd["exception_pairs"] = pairs
    sysd["exception_sigmas"] = np.array([2.5, 3.0, 2.0])
    sysd["exception_epsilons"] = np.array([0.1, 0.05, 0.2])
    sysd["exception_chargeprods"] = np.array([0.1, -0.2, 0.05])
    sysd = f32_round(sysd)
    _, _, _, _, _, info = _compare("harmonic_imp_exc_ub", sysd)
    assert info["total_rel"] < E_TOL
    assert info["force_rel_rms"] < F_TOL


@pytest.mark.parametrize("theta_deg", [179.0, 179.6, 179.85])
def test_near_straight_angle_and_collinear_torsions(theta_deg):
    """The 8-atom chain (golden chain8) with its angle 1-2-3 bent to nearly
    180 degrees (inside the reference's cos clip at +-0.999999) and placed
    35 A from the origin: the angle's force and those of the two torsions
    through the straight triple are ill-conditioned in float32 (1 / sin theta,
    |b1 x b2| << |b1||b2|) and must still match float64 to the force bar,
    atom by atom."""
    from tests.golden import make_golden as G

    sysd = G.chain8()
    x = np.asarray(sysd["positions"], np.float64) + 35.0
    u = (x[1] - x[2]) / np.linalg.norm(x[1] - x[2])
    w = np.cross(u, [0.3, 0.4, 0.866])
    w /= np.linalg.norm(w)
    th = np.radians(theta_deg)
    shift = x[2] + 1.52 * (np.cos(th) * u + np.sin(th) * w) - x[3]
    x[[3, 4, 5, 7]] += shift  # atom 3 and the atoms behind it move together
    sysd["positions"] = x
    sysd = f32_round(sysd)
    xs = sysd["positions"]
    v1, v2 = xs[1] - xs[2], xs[3] - xs[2]
    c = v1 @ v2 / (np.linalg.norm(v1) * np.linalg.norm(v2))
    assert -0.999999 < c < -0.99984  # inside the clip, within 1 degree of straight
    _, f, _, terms, fo, info = _compare(f"straight_angle_{theta_deg}", sysd)
    assert info["total_rel"] < E_TOL
    assert info["force_rel_rms"] < F_TOL
    assert np.max(np.abs(f - fo)) < F_TOL * np.max(np.abs(fo))


def test_deterministic_and_batch_equals_single():
    """Fixed-order accumulation: repeated calls and batched vs alone are bitwise equal."""
    from prolix_amd.engine import GBEngine

    systems = [protein(n, 20 + k, radius=synth.compact_radius(n)) for k, n in enumerate((97, 300, 513))]
    print("[gpu] batch engine", flush=True)
    eng = GBEngine(systems)
    pos = eng.pack([s["positions"] for s in systems])
    e1, f1, _ = eng.energy_forces(pos)
    e2, f2, _ = eng.energy_forces(pos)
    assert np.array_equal(f1.cpu().numpy(), f2.cpu().numpy())
    assert np.array_equal(e1.cpu().numpy(), e2.cpu().numpy())
    fb = eng.unpack(f1)
    eb = e1.cpu().numpy()
    eng.close()
    for k, s in enumerate(systems):
        es, fs, _ = _gpu_eval(s)
        assert np.array_equal(fs, fb[k]), k
        assert np.array_equal(es, eb[k]), k
