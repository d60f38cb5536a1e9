"""ctypes wrapper of oracle/cpu_engine.cpp (built as oracle/liboracle_cpu.so).

TEST INFRASTRUCTURE ONLY: the large-N float64 checker and the timed CPU
baseline. Only tests, the smoke check and the benchmark's CPU-baseline leg use it.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

from oracle import gb_oracle as O

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle_cpu.so"
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(os.fspath(LIB))
        L.oracle_cpu_create.restype = ctypes.c_void_p
        L.oracle_cpu_create.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_cpu_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_cpu_set_f32_cutoff.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.oracle_cpu_threads.restype = ctypes.c_int
        for name in ("oracle_cpu_energy_forces_f64", "oracle_cpu_energy_forces_f32"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        for name in ("oracle_cpu_baoab_f64", "oracle_cpu_baoab_f32"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_uint64, ctypes.c_uint32]
        _lib = L
    return _lib


class CpuSystem:
    """One system on the CPU restatement (float64 or float32 arithmetic)."""

    def __init__(self, sysd: dict, f32_cutoff: bool = False, **opts):
        """``f32_cutoff``: classify pairs against the cutoff with the float32
        r^2 test of the HIP kernels (the arithmetic stays float64)."""
        from prolix_amd import _native as N  # ABI struct layout only (data marshalling)

        self._desc = N.Desc(sysd, **opts)
        coeff = None
        if sysd.get("cmap_torsions") is not None and np.size(sysd.get("cmap_torsions")):
            coeff = np.ascontiguousarray(O.cmap_coefficients(sysd["cmap_energy_grids"]), np.float64)
        self._coeff = coeff
        self._h = lib().oracle_cpu_create(ctypes.byref(self._desc.desc),
                                          None if coeff is None else coeff.ctypes.data_as(ctypes.c_void_p))
        self.n = self._desc.n_atoms
        lib().oracle_cpu_set_f32_cutoff(self._h, 1 if f32_cutoff else 0)

    def energy_forces(self, pos, dtype=np.float64):
        x = np.ascontiguousarray(pos, dtype=dtype).reshape(-1, 3)
        f = np.zeros_like(x)
        e = np.zeros(12)
        fn = lib().oracle_cpu_energy_forces_f64 if dtype == np.float64 else lib().oracle_cpu_energy_forces_f32
        fn(self._h, x.ctypes.data, f.ctypes.data, e.ctypes.data)
        return e, f

    def baoab(self, pos, mom, force, steps, dt_fs, temperature, gamma_ps, seed, step0=0, dtype=np.float64):
        fn = lib().oracle_cpu_baoab_f64 if dtype == np.float64 else lib().oracle_cpu_baoab_f32
        fn(self._h, pos.ctypes.data, mom.ctypes.data, force.ctypes.data, int(steps), float(dt_fs),
           float(temperature), float(gamma_ps), int(seed), int(step0))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_cpu_destroy(self._h)
            self._h = None


def threads() -> int:
    return int(lib().oracle_cpu_threads())
