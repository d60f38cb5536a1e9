"""The persistent step kernel (prx_persist.hip) against the captured-graph
path: the same per-item and per-tile device code behind grid-wide barriers,
so trajectories agree bit for bit -- one small system, one of config-1 size,
several packed systems, and a neighbour-list system. The wall time per step
of both paths is printed."""

from __future__ import annotations

import time

import numpy as np
import pytest

from prolix_amd import synth
from tests.helpers import protein

pytestmark = pytest.mark.gpu


def _run(systems, mode, steps, **kw):
    import torch

    from prolix_amd.engine import GBEngine, Langevin

    eng = GBEngine(systems, **kw)
    eng.set_step_mode(mode)
    n = eng.n_systems
    eng.init([s["positions"] for s in systems], [Langevin(1.0, 300.0, 1.0, seed=50 + k, safety_caps=True)
                                                 for k in range(n)])
    eng.steps(5)
    torch.cuda.synchronize()
    t = time.perf_counter()
    eng.steps(steps)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t) / steps * 1e3
    out = [a.cpu().numpy() for a in eng.state()]
    counts = [eng.step_count(k) for k in range(n)]
    eng.close()
    return out, counts, ms


@pytest.mark.parametrize("case", ["peptide138", "globule2500", "three_packed", "list_2000"])
def test_persistent_equals_graph(case):
    kw = {}
    if case == "peptide138":
        systems = [protein(138, 0)]
    elif case == "globule2500":
        systems = [protein(2500, 1, radius=synth.compact_radius(2500))]
    elif case == "three_packed":
        systems = [protein(int(n), 70 + k, radius=synth.compact_radius(int(n))) for k, n in enumerate((300, 900, 1700))]
    else:
        systems = [protein(2000, 9, radius=synth.compact_radius(2000))]
        kw = dict(cutoff=10.0, gb_cutoff=True)
    print(f"[gpu] persistent vs graph: {case}", flush=True)
    g, cg, ms_g = _run(systems, "graph", 60, **kw)
    p, cp, ms_p = _run(systems, "persistent", 60, **kw)
    print(f"[persist] {case}: graph {ms_g:.4f} ms/step, persistent {ms_p:.4f} ms/step", flush=True)
    assert cg == cp == [65] * len(systems)
    for a, b in zip(g, p):
        assert np.array_equal(a, b)
