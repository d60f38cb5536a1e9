"""Cutoff paths on the device: golden vectors, the neighbour-list GB form
(Born sums and GB pairs over the pairs closer than the cutoff), the tile
classes (near / straddle / far) and mixed cutoff modes in one engine.

Tolerances are those of the dense parity tests: energy relative error
<= 1e-6 (relative to max(|E_total|, 0.2 * sum_k |E_k|)) and force RMS
relative error <= 1e-5 against float64 on the same float32-rounded inputs.
A hard cutoff makes the energy jump where a pair crosses it, and at a few
thousand atoms some pairs lie within float32 rounding of any cutoff. The
large cases are therefore checked against the float64 CPU restatement with
its cutoff test in the kernels' float32 form (``f32_cutoff``, pinned on the
CPU against an exact-rational emulation of that test), so both sides sum
the same set of pairs; the golden vectors keep every pair >= 5e-5 A from
the cutoff and compare with the plain float64 oracle.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import gb_oracle as O
from prolix_amd import _native as N
from prolix_amd import synth
from tests.golden import make_golden as G
from tests.helpers import protein, rel_rms

pytestmark = pytest.mark.gpu

E_TOL = 1e-6
F_TOL = 1e-5


def _gpu(sysd, **kw):
    from prolix_amd.engine import GBEngine

    print(f"[gpu] energy_forces n={len(sysd['charges'])} {kw}", flush=True)
    eng = GBEngine(sysd, **kw)
    e, f, b = eng.energy_forces(sysd["positions"], energies=True, born_radii=True)
    out = (e.cpu().numpy()[0], f.cpu().numpy(), b.cpu().numpy())
    eng.close()
    return out


def _e_scale(terms):
    return max(abs(terms[-1]), 0.2 * float(np.sum(np.abs(terms[:-1]))), 1e-30)


@pytest.mark.parametrize("name", ("peptide138", "peptide138_cut9", "chain8", "chain8_harmonic", "ions3",
                                  "peptide138_gbcut12", "chain8_gbcut4"))
def test_golden_vectors(name):
    sysd, cutoff, gbc, terms, forces, born = G.load(name)
    e, f, b = _gpu(sysd, cutoff=cutoff, gb_cutoff=gbc)
    rel = abs(e[-1] - terms[-1]) / _e_scale(terms)
    fr = rel_rms(f, forces)
    print(f"[{name}] total rel {rel:.2e} force rel-rms {fr:.2e}", flush=True)
    assert rel < E_TOL
    assert np.all(np.abs(e[:-1] - terms[:-1]) <= E_TOL * np.sum(np.abs(terms[:-1])) + 1e-9)
    assert fr < F_TOL
    assert np.max(np.abs(b - born) / born) < 1e-5


def test_gb_cutoff_globule_2500():
    """Config-1-sized globule with the neighbour-list GB form at ~12 A: far,
    straddling and near items all occur."""
    from oracle import cpu as C

    s = protein(2500, 1, radius=synth.compact_radius(2500))
    rc = 12.0
    e, f, b = _gpu(s, cutoff=rc, gb_cutoff=True)
    eo, fo = C.CpuSystem(s, cutoff=rc, gb_cutoff=True, f32_cutoff=True).energy_forces(s["positions"])
    rel = abs(e[-1] - eo[-1]) / _e_scale(eo)
    fr = rel_rms(f, fo)
    print(f"[gbcut_2500] total rel {rel:.2e} force rel-rms {fr:.2e}", flush=True)
    assert rel < E_TOL
    assert fr < F_TOL
    assert np.all(np.abs(e[:-1] - eo[:-1]) <= E_TOL * np.sum(np.abs(eo[:-1])) + 1e-9)


def test_lj_coulomb_cutoff_globule_2500():
    """Drop-in semantics: 9 A cutoff on LJ/Coulomb, GB over every pair (far
    items run GB only)."""
    from oracle import cpu as C

    s = protein(2500, 2, radius=synth.compact_radius(2500))
    rc = 9.0
    e, f, _ = _gpu(s, cutoff=rc)
    eo, fo = C.CpuSystem(s, cutoff=rc, f32_cutoff=True).energy_forces(s["positions"])
    rel = abs(e[-1] - eo[-1]) / _e_scale(eo)
    fr = rel_rms(f, fo)
    print(f"[cut9_2500] total rel {rel:.2e} force rel-rms {fr:.2e}", flush=True)
    assert rel < E_TOL
    assert fr < F_TOL


def test_complete_list_equals_dense():
    """Cutoff beyond the largest distance: every item is near, and the
    neighbour-list form gives the dense result."""
    s = protein(700, 3, radius=synth.compact_radius(700))
    x = s["positions"]
    rmax = float(np.max(np.linalg.norm(x[:, None] - x[None], axis=-1)))
    e0, f0, b0 = _gpu(s)
    e1, f1, b1 = _gpu(s, cutoff=rmax + 5.0, gb_cutoff=True)
    assert abs(e1[-1] - e0[-1]) <= 1e-7 * _e_scale(e0)
    assert rel_rms(f1, f0) < 1e-7
    assert np.max(np.abs(b1 - b0) / b0) < 1e-7


def test_mixed_cutoff_modes_batch_equals_alone():
    """Systems with different cutoff modes share one pool; each result equals the
    system alone in a pool of its own, bitwise. (Pools keep no neighbour lists:
    a gb_cut system sweeps its tile pairs there; an engine's list evaluation
    agrees with that to rounding, tests/test_gpu_nl.py. The stored forces come
    from the force-only variant of the pair kernel, energies from the energy
    variant; the two are compiled separately and may round differently.)"""
    from prolix_amd.engine import GBEngine, GBPool, Langevin

    systems = [protein(n, 40 + k, radius=synth.compact_radius(n)) for k, n in enumerate((900, 600, 1300))]
    modes = [dict(), dict(cutoff=11.0, gb_cutoff=True), dict(cutoff=9.0)]
    print("[gpu] mixed-mode pool", flush=True)
    pool = GBPool(max_atoms=sum(len(s["charges"]) for s in systems) + 64, max_systems=4)
    slots = [pool.admit(s, Langevin(1.0, 300.0, 1.0, seed=k), **m) for k, (s, m) in enumerate(zip(systems, modes))]
    ens = [pool.energies(sl) for sl in slots]  # an energy evaluation of the whole pool
    got = [pool.read(sl) for sl in slots]
    pool.close()
    for k, (s, m) in enumerate(zip(systems, modes)):
        one = GBPool(max_atoms=len(s["charges"]), max_systems=1)
        sl = one.admit(s, Langevin(1.0, 300.0, 1.0, seed=k), **m)
        e = one.energies(sl)
        f = one.read(sl)[2]
        one.close()
        assert np.array_equal(got[k][2], f), k
        assert np.array_equal(ens[k], e), k
        if not m.get("gb_cutoff"):  # without a list, an engine runs the same sweep
            eng = GBEngine(s, **m)
            e2, _, _ = eng.energy_forces(s["positions"], energies=True)
            _, f2, _ = eng.energy_forces(s["positions"], energies=False)
            eng.close()
            assert np.array_equal(got[k][2], f2.cpu().numpy()), k
            assert np.array_equal(ens[k], e2.cpu().numpy()[0]), k


def test_large_gb_cutoff_20k():
    """20,000-atom globule, 20 A cutoff with the neighbour-list GB form, against
    the float64 CPU restatement."""
    from oracle import cpu as C

    s = protein(20000, 7, radius=synth.compact_radius(20000))
    rc = 20.0
    e, f, _ = _gpu(s, cutoff=rc, gb_cutoff=True)
    eo, fo = C.CpuSystem(s, cutoff=rc, gb_cutoff=True, f32_cutoff=True).energy_forces(s["positions"])
    rel = abs(e[-1] - eo[-1]) / _e_scale(eo)
    fr = rel_rms(f, fo)
    print(f"[gbcut_20k] total rel {rel:.2e} force rel-rms {fr:.2e}", flush=True)
    assert rel < E_TOL
    assert fr < F_TOL


@pytest.mark.parametrize("gb_cutoff", (False, True))
def test_boundary_pair_follows_float32_test(gb_cutoff):
    """A pair whose exact length is above the cutoff but whose float32 r^2 (as
    the kernels form it) is below: the device counts it, like the
    restatement's float32 classification."""
    from tests.helpers import boundary_pair

    rc = 12.0
    d = boundary_pair(rc)
    s = {"positions": np.array([[0.0, 0.0, 0.0], d.astype(np.float64)]), "charges": np.array([1.0, -1.0]),
         "sigmas": np.array([2.0, 2.0]), "epsilons": np.array([0.1, 0.1]), "radii": np.array([1.5, 1.5])}
    e, _, b = _gpu(s, cutoff=rc, gb_cutoff=gb_cutoff)
    assert e[N.TERMS.index("coulomb")] == pytest.approx(-332.0637 / rc, rel=1e-5)
    if gb_cutoff:  # the partner descreens: B above the isolated-ion value
        assert np.all(b > 1.41 * (1 + 1e-6))
