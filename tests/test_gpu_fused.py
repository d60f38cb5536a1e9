"""The fused step kernel (prx_fused.hip: one workgroup per system, workgroup
barriers only) against the captured-graph path: the same per-item and
per-tile device code on the same work lists, so trajectories agree bit for
bit -- one small system, several packed systems of different sizes, a
1,000-atom system, and a pool whose systems stop at their step limits. The
wall time per step of both paths is printed; the auto mode runs the fused
kernel up to 192 atoms per system and graphs above."""

from __future__ import annotations

import time

import numpy as np
import pytest

from prolix_amd import synth
from tests.helpers import protein

pytestmark = pytest.mark.gpu


def _run(systems, mode, steps, warm=5):
    import torch

    from prolix_amd.engine import GBEngine, Langevin

    eng = GBEngine(systems)
    eng.set_step_mode(mode)
    n = eng.n_systems
    eng.init([s["positions"] for s in systems], [Langevin(1.0, 300.0, 1.0, seed=80 + k, safety_caps=True)
                                                 for k in range(n)])
    eng.steps(warm)
    torch.cuda.synchronize()
    t = time.perf_counter()
    eng.steps(steps)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t) / steps * 1e3
    out = [a.cpu().numpy() for a in eng.state()]
    counts = [eng.step_count(k) for k in range(n)]
    eng.close()
    return out, counts, ms


@pytest.mark.parametrize("case", ["peptide138", "small_packed", "atoms1000"])
def test_fused_equals_graph(case):
    if case == "peptide138":
        systems = [protein(138, 0)]
    elif case == "small_packed":
        systems = [protein(int(n), 90 + k, radius=synth.compact_radius(int(n))) for k, n in enumerate((138, 250, 380))]
    else:
        systems = [protein(1000, 5, radius=synth.compact_radius(1000))]
    print(f"[gpu] fused vs graph: {case}", flush=True)
    g, cg, ms_g = _run(systems, "graph", 200)
    f, cf, ms_f = _run(systems, "fused", 200)
    print(f"[fused] {case}: graph {ms_g:.4f} ms/step, fused {ms_f:.4f} ms/step", flush=True)
    assert cg == cf == [205] * len(systems)
    for a, b in zip(g, f):
        assert np.array_equal(a, b)


def test_fused_step_times_by_size():
    """Graph and fused ms/step for one system of 138 to 1,000 atoms (printed)
    beside the auto mode's."""
    from prolix_amd.engine import GBEngine

    for n in (138, 256, 384, 512, 640, 1000):
        s = protein(n, 40 + n % 7, radius=synth.compact_radius(n))
        _, _, ms_g = _run([s], "graph", 300)
        _, _, ms_f = _run([s], "fused", 300)
        _, _, ms_a = _run([s], "auto", 300)
        eng = GBEngine(s)
        items = eng.work_items()
        eng.close()
        print(f"[fused] {n} atoms ({items} items): graph {ms_g:.4f}, fused {ms_f:.4f}, auto {ms_a:.4f} ms/step",
              flush=True)


def test_fused_pool_with_step_limits():
    """A pool of small systems runs fused; a system at its step limit stops
    there exactly as on the graph path."""
    from prolix_amd.engine import GBPool, Langevin

    systems = [protein(n, 170 + k, radius=synth.compact_radius(n)) for k, n in enumerate((150, 300))]
    params = [Langevin(1.0, 300.0, 1.0, seed=710 + k, safety_caps=True) for k in range(2)]
    out = {}
    for mode in ("graph", "fused"):
        print(f"[gpu] pool step limits, {mode}", flush=True)
        pool = GBPool(max_atoms=1000, max_systems=4)
        pool.set_step_mode(mode)
        slots = [pool.admit(s, p) for s, p in zip(systems, params)]
        pool.set_step_limit(slots[0], 60)
        pool.steps(100)
        out[mode] = [pool.read(sl) for sl in slots]
        pool.close()
    for a, b in zip(out["graph"], out["fused"]):
        assert a[3] == b[3]
        for x, y in zip(a[:3], b[:3]):
            assert np.array_equal(x, y)
    assert out["fused"][0][3] == 60 and out["fused"][1][3] == 100
