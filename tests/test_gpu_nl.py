"""The device neighbour list of gb_cut systems (prx_nl.hip): per 64-atom tile
the atoms of later tiles within cutoff + skin of its box, evaluated in
chunks of 64 with the per-pair cutoff test.

- the list engine equals the all-tile-pairs sweep of a serving pool (which
  keeps no lists) to float32 rounding, and the float64 restatement to the
  parity tolerances;
- with a Verlet skin the list is rebuilt only when an atom moved more than
  skin / 2 (jax_md's rule, physics/neighbor_list.py:155-172): after steps
  that reused lists, the forces the step graph stored equal a fresh list's
  to rounding, and the float64 restatement at those positions;
- several list systems and a dense one packed in one engine give each
  system's own result bit for bit (lists never cross systems).
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import cpu as C
from prolix_amd import synth
from tests.helpers import protein, rel_rms

pytestmark = pytest.mark.gpu

E_TOL = 1e-6
F_TOL = 1e-5


def _e_scale(terms):
    return max(abs(terms[-1]), 0.2 * float(np.sum(np.abs(terms[:-1]))), 1e-30)


def _oracle(s, x, rc):
    return C.CpuSystem(s, cutoff=rc, gb_cutoff=True, f32_cutoff=True).energy_forces(np.asarray(x, np.float64))


def test_list_matches_all_tile_pairs_and_float64():
    from prolix_amd.engine import GBEngine, GBPool, Langevin

    s = protein(3000, 11, radius=synth.compact_radius(3000))
    rc = 12.0
    print("[gpu] list engine vs pool sweep, 3000 atoms, 12 A", flush=True)
    eng = GBEngine(s, cutoff=rc, gb_cutoff=True)
    e, f, _ = eng.energy_forces(s["positions"], energies=True)
    e, f = e.cpu().numpy()[0], f.cpu().numpy()
    eng.close()
    pool = GBPool(max_atoms=4000, max_systems=4)
    slot = pool.admit(s, Langevin(1.0, 300.0, 1.0, seed=1), cutoff=rc, gb_cutoff=True)
    ep = pool.energies(slot)
    _, _, fp, _ = pool.read(slot)
    pool.close()
    eo, fo = _oracle(s, s["positions"], rc)
    print(f"[nl] list vs sweep: energy rel {abs(e[-1] - ep[-1]) / abs(ep[-1]):.2e}, force {rel_rms(f, fp):.2e}; "
          f"vs float64: energy rel {abs(e[-1] - eo[-1]) / _e_scale(eo):.2e}, force {rel_rms(f, fo):.2e}", flush=True)
    assert abs(e[-1] - ep[-1]) <= 1e-6 * _e_scale(ep)
    assert rel_rms(f, fp) < 1e-6
    assert abs(e[-1] - eo[-1]) / _e_scale(eo) < E_TOL
    assert rel_rms(f, fo) < F_TOL
    assert np.all(np.abs(e[:-1] - eo[:-1]) <= E_TOL * np.sum(np.abs(eo[:-1])) + 1e-9)


@pytest.mark.parametrize("skin", [0.0, 2.0])
def test_skin_reuse_keeps_forces_exact(skin):
    from prolix_amd.engine import GBEngine, Langevin

    s = protein(2500, 12, radius=synth.compact_radius(2500))
    rc = 12.0
    print(f"[gpu] list skin {skin}, 2500 atoms, 40 steps", flush=True)
    eng = GBEngine(s, cutoff=rc, gb_cutoff=True)
    eng.set_nl_skin(skin)
    eng.init(s["positions"], Langevin(1.0, 300.0, 1.0, seed=4, safety_caps=True))
    eng.steps(40)
    x, _, f_step = eng.state()
    e1, f1, _ = eng.energy_forces(x, energies=True)  # a fresh list at the final positions
    x, f_step, f1, e1 = x.cpu().numpy(), f_step.cpu().numpy(), f1.cpu().numpy(), e1.cpu().numpy()[0]
    warn = eng.warnings(0)
    eng.close()
    eo, fo = _oracle(s, x, rc)
    print(f"[nl] skin {skin}: stored vs fresh force {rel_rms(f_step, f1):.2e}; vs float64 energy "
          f"{abs(e1[-1] - eo[-1]) / _e_scale(eo):.2e} force {rel_rms(f1, fo):.2e}; caps {warn}", flush=True)
    if warn[1] == 0:
        assert rel_rms(f_step, f1) < 1e-6
    assert abs(e1[-1] - eo[-1]) / _e_scale(eo) < E_TOL
    assert rel_rms(f1, fo) < F_TOL


def test_list_systems_packed_equal_alone():
    from prolix_amd.engine import GBEngine, Langevin

    rc = 10.0
    systems = [protein(int(n), 30 + k, radius=synth.compact_radius(int(n))) for k, n in enumerate((900, 1500, 2100))]
    print("[gpu] 3 list systems packed vs alone", flush=True)
    eng = GBEngine(systems, cutoff=rc, gb_cutoff=True)
    # the batched step's caps: the unrelaxed compact starts stay bounded
    eng.init([s["positions"] for s in systems], [Langevin(1.0, 300.0, 1.0, seed=7 + k, safety_caps=True)
                                                 for k in range(3)])
    eng.steps(25)
    x, p, f = (a.cpu().numpy() for a in eng.state())
    off = eng.offsets
    eng.close()
    for k, s in enumerate(systems):
        one = GBEngine(s, cutoff=rc, gb_cutoff=True)
        one.init(s["positions"], Langevin(1.0, 300.0, 1.0, seed=7 + k, safety_caps=True))
        one.steps(25)
        xa, pa, fa = (a.cpu().numpy() for a in one.state())
        one.close()
        sl = slice(off[k], off[k + 1])
        assert np.isfinite(xa).all() and np.isfinite(pa).all()
        assert np.array_equal(xa, x[sl]) and np.array_equal(pa, p[sl]) and np.array_equal(fa, f[sl])


def test_dense_neighbor_list_object():
    """make_neighbor_list_fn (neighbor_list.py:155-172): every pair within
    cutoff + 0.5 A, both directions, no self pairs, rows in increasing index,
    the atom count as the sentinel; update keeps the list below half the skin
    and rebuilds above it; a list too small reports the overflow."""
    import torch
    from scipy.spatial import cKDTree

    from prolix_amd.physics.neighbor_list import make_neighbor_list_fn

    s = protein(2500, 13, radius=synth.compact_radius(2500))
    x = np.asarray(s["positions"], np.float32)
    n = len(x)
    rc = 9.0
    print("[gpu] dense neighbour list, 2500 atoms, 9 + 0.5 A", flush=True)
    fns = make_neighbor_list_fn(None, None, rc)
    nb = fns.allocate(x)
    idx = nb.idx.cpu().numpy()
    assert not nb.did_buffer_overflow and idx.shape[1] >= nb.max_occupancy
    # the exact float32 test of the kernel, away from the boundary
    x64 = x.astype(np.float64)
    r = rc + 0.5
    tree = cKDTree(x64)
    sure = tree.query_ball_point(x64, r - 1e-4)
    maybe = tree.query_ball_point(x64, r + 1e-4)
    for i in range(n):
        row = idx[i][idx[i] < n]
        assert np.all(np.diff(row) > 0) and i not in row and np.all(idx[i][len(row):] == n)
        got = set(row.tolist())
        assert set(sure[i]) - {i} <= got <= set(maybe[i])
    same = fns.update(x + 0.1, nb)  # every atom moved 0.17 A < dr_threshold / 2 = 0.25 A: kept
    assert same is nb
    moved = x.copy()
    moved[7] += 0.3
    nb2 = fns.update(moved, nb)
    assert nb2 is not nb and nb2.idx.shape == nb.idx.shape
    small = fns._build(x, 8)  # capacity 8: overflow reported
    assert small.did_buffer_overflow and small.max_occupancy == nb.max_occupancy
    torch.cuda.synchronize()
