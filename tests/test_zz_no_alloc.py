"""prx_energy_forces_noalloc (the allocation-free one-shot evaluation) against
prx_energy_forces on an engine, through the C ABI: two calls in a row, the
second with a smaller system, each equal to the engine's result bit for bit
(same kernels on the same inputs). Also: the size limits are rejected with
an error, and the call allocates no device memory (the free device memory
reads the same before and after)."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from prolix_amd import _native as N
from prolix_amd import synth
from tests.helpers import protein

pytestmark = pytest.mark.gpu


def _vp(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _noalloc(torch, sysd):
    """Returns (status, free device bytes before, after, energies, forces, Born radii)."""
    d = N.Desc(sysd)
    n = d.n_atoms
    x = torch.as_tensor(np.asarray(sysd["positions"], np.float32), device="cuda").contiguous()
    f = torch.empty((n, 3), dtype=torch.float32, device="cuda")
    e = torch.empty((1, N.N_ENERGY), dtype=torch.float64, device="cuda")
    b = torch.empty((n,), dtype=torch.float32, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    code = N.lib().prx_energy_forces_noalloc(ctypes.byref(d.desc), _vp(x), _vp(f), _vp(e), _vp(b), stream)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    return code, free0, free1, e, f, b


def _engine(sysd):
    from prolix_amd.engine import GBEngine

    eng = GBEngine(sysd)
    e, f, b = eng.energy_forces(sysd["positions"], energies=True, born_radii=True)
    out = (e.cpu().numpy(), f.cpu().numpy(), b.cpu().numpy())
    eng.close()
    return out


def test_noalloc_matches_engine_on_two_calls_in_a_row():
    import torch

    big = protein(2500, 1, radius=synth.compact_radius(2500))
    small = protein(138, 0)
    for label, sysd in (("2500-atom globule", big), ("138-atom peptide (second, smaller call)", small)):
        print(f"[gpu] noalloc vs engine: {label}", flush=True)
        ref_e, ref_f, ref_b = _engine(sysd)
        code, free0, free1, e, f, b = _noalloc(torch, sysd)
        assert code == 0, N.lib().prx_last_error().decode()
        assert free1 == free0  # the call allocated no device memory
        assert np.array_equal(e.cpu().numpy(), ref_e)
        assert np.array_equal(f.cpu().numpy(), ref_f)
        assert np.array_equal(b.cpu().numpy(), ref_b)


def test_noalloc_rejects_systems_beyond_its_limits():
    import torch

    print("[gpu] noalloc limits", flush=True)
    n = 8193
    sysd = {"positions": np.zeros((n, 3)), "charges": np.zeros(n), "sigmas": np.ones(n), "epsilons": np.zeros(n),
            "radii": np.full(n, 1.5)}
    d = N.Desc(sysd)
    x = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
    f = torch.empty_like(x)
    code = N.lib().prx_energy_forces_noalloc(ctypes.byref(d.desc), _vp(x), _vp(f), None, None,
                                             ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert code == 1 and b"8192" in N.lib().prx_last_error()
