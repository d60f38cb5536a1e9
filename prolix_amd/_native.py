"""ctypes binding of the engine's C ABI (include/prolix_gb.h).

The library is loaded from ``prolix_amd/lib/libprolix_gb.so`` (built in-tree by
``prolix_amd.build``). There is no fallback: if the library is missing or
fails to load, every entry point raises.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libprolix_gb.so"

PRX_OK = 0
ERRORS = {1: "invalid argument", 2: "HIP runtime error", 3: "capacity exhausted", 4: "call out of order"}

TERMS = ("bond", "angle", "urey_bradley", "dihedral", "improper", "cmap", "lj", "exception", "gb",
         "coulomb", "ace", "total")
N_ENERGY = len(TERMS)

c_i32p = ctypes.POINTER(ctypes.c_int32)
c_f64p = ctypes.POINTER(ctypes.c_double)
c_f32p = ctypes.POINTER(ctypes.c_float)


class SystemDesc(ctypes.Structure):
    _fields_ = [
        ("n_atoms", ctypes.c_int32),
        ("charges", c_f64p), ("sigmas", c_f64p), ("epsilons", c_f64p), ("radii", c_f64p),
        ("scaled_radii", c_f64p), ("masses", c_f64p),
        ("n_bonds", ctypes.c_int32), ("bond_idx", c_i32p), ("bond_params", c_f64p),
        ("n_ub", ctypes.c_int32), ("ub_idx", c_i32p), ("ub_params", c_f64p),
        ("n_angles", ctypes.c_int32), ("angle_idx", c_i32p), ("angle_params", c_f64p),
        ("n_dihedrals", ctypes.c_int32), ("dihedral_terms", ctypes.c_int32), ("dihedral_idx", c_i32p),
        ("dihedral_params", c_f64p),
        ("n_impropers", ctypes.c_int32), ("improper_terms", ctypes.c_int32),
        ("improper_param_width", ctypes.c_int32), ("improper_idx", c_i32p), ("improper_params", c_f64p),
        ("n_cmap", ctypes.c_int32), ("cmap_torsions", c_i32p), ("cmap_map", c_i32p),
        ("n_cmap_maps", ctypes.c_int32), ("cmap_grid", ctypes.c_int32), ("cmap_grids", c_f64p),
        ("excl_cols", ctypes.c_int32), ("excl_idx", c_i32p), ("excl_scale_vdw", c_f64p),
        ("excl_scale_elec", c_f64p),
        ("n_exceptions", ctypes.c_int32), ("exc_pairs", c_i32p), ("exc_sigma", c_f64p),
        ("exc_epsilon", c_f64p), ("exc_chargeprod", c_f64p),
        ("cutoff", ctypes.c_double), ("coulomb_constant", ctypes.c_double),
        ("gb_coulomb_constant", ctypes.c_double), ("solute_dielectric", ctypes.c_double),
        ("solvent_dielectric", ctypes.c_double), ("dielectric_offset", ctypes.c_double),
        ("implicit_solvent", ctypes.c_int32), ("gb_cutoff", ctypes.c_int32),
    ]


class LangevinParams(ctypes.Structure):
    _fields_ = [("dt_fs", ctypes.c_double), ("temperature", ctypes.c_double), ("gamma_ps", ctypes.c_double),
                ("seed", ctypes.c_uint64), ("safety_caps", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class PoolLimits(ctypes.Structure):
    _fields_ = [("max_atoms", ctypes.c_int64), ("max_bonded", ctypes.c_int64), ("max_cmap_coeff", ctypes.c_int64),
                ("max_excl_tiles", ctypes.c_int64), ("max_systems", ctypes.c_int32)]


_SIGS = {
    "prx_last_error": (ctypes.c_char_p, []),
    "prx_abi_version": (ctypes.c_int32, []),
    "prx_system_create": (ctypes.c_int, [ctypes.POINTER(SystemDesc), ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    "prx_engine_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "prx_engine_layout": (ctypes.c_int, [ctypes.c_void_p, c_i32p, ctypes.c_int32]),
    "prx_energy_forces": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p]),
    "prx_energy_forces_noalloc": (ctypes.c_int, [ctypes.POINTER(SystemDesc), ctypes.c_void_p, ctypes.c_void_p,
                                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "prx_integrator_init": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.POINTER(LangevinParams), ctypes.c_void_p]),
    "prx_integrator_steps": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]),
    "prx_engine_set_masses": (ctypes.c_int, [ctypes.c_void_p, c_f64p]),
    "prx_integrator_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p]),
    "prx_integrator_kinetic": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "prx_recenter_com": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "prx_integrator_step_count": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)]),
    "prx_integrator_warnings": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)]),
    "prx_integrator_set_step_counts": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
                                                      ctypes.c_void_p]),
    "prx_integrator_health": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p]),
    "prx_engine_set_fork_min_atoms": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    "prx_engine_set_nl_skin": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "prx_engine_set_step_mode": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "prx_pool_create": (ctypes.c_int, [ctypes.POINTER(PoolLimits), ctypes.POINTER(ctypes.c_void_p)]),
    "prx_pool_admit": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SystemDesc), c_f32p,
                                      ctypes.POINTER(LangevinParams), c_i32p, ctypes.c_void_p]),
    "prx_pool_release": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]),
    "prx_pool_set_step_limit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p]),
    "prx_pool_read": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, c_f32p, c_f32p, c_f32p,
                                     ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p]),
    "prx_pool_slots": (ctypes.c_int, [ctypes.c_void_p, c_i32p, ctypes.c_int32, c_i32p]),
    "prx_pool_energy": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, c_f64p, ctypes.c_void_p]),
    "prx_profile_passes": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, c_f64p]),
    "prx_profile_steps": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, c_f64p]),
    "prx_engine_work_items": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "prx_neighbor_list_dense": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, ctypes.c_int32,
                                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "prx_cmap_coefficients": (ctypes.c_int, [c_f64p, ctypes.c_int32, c_f64p]),
    "prx_plan_tiles": (ctypes.c_int, [ctypes.POINTER(SystemDesc), ctypes.POINTER(ctypes.c_int64),
                                      ctypes.POINTER(ctypes.c_int64)]),
}

_lib = None


class NativeError(RuntimeError):
    pass


def lib() -> ctypes.CDLL:
    """The loaded engine library; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise NativeError(f"{LIB_PATH} is missing: build it with `python -m prolix_amd.build` "
                          "(there is no non-native fallback)")
    L = ctypes.CDLL(os.fspath(LIB_PATH))
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(code: int, what: str) -> None:
    if code != PRX_OK:
        msg = lib().prx_last_error().decode(errors="replace")
        raise NativeError(f"{what}: {ERRORS.get(code, code)}: {msg}")


def _arr(x, dtype, shape_tail=None):
    if x is None:
        return None
    a = np.ascontiguousarray(np.asarray(x), dtype=dtype)
    if shape_tail is not None:
        a = a.reshape((-1,) + tuple(shape_tail))
    return a


def _ptr(a, ctype):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctype))


class Desc:
    """A SystemDesc plus the numpy arrays it points into (kept alive together)."""

    def __init__(self, sysd: dict, *, cutoff: float = 0.0, coulomb_constant: float = 332.0637,
                 gb_coulomb_constant: float = 332.0637, solute_dielectric: float = 1.0,
                 solvent_dielectric: float = 78.5, dielectric_offset: float = 0.09,
                 implicit_solvent: bool = True, gb_cutoff: bool = False):
        keep = {}

        def f64(name, tail=None):
            a = _arr(sysd.get(name), np.float64, tail)
            keep[name] = a
            return a

        def i32(name, tail=None):
            a = _arr(sysd.get(name), np.int32, tail)
            keep[name] = a
            return a

        d = SystemDesc()
        q = f64("charges")
        if q is None or q.size == 0:
            raise ValueError("system['charges'] is empty -- cannot infer n_atoms")
        n = int(q.shape[0])
        d.n_atoms = n
        d.charges = _ptr(q, ctypes.c_double)
        d.sigmas = _ptr(f64("sigmas"), ctypes.c_double)
        d.epsilons = _ptr(f64("epsilons"), ctypes.c_double)
        radii = f64("radii")
        if radii is None:
            radii = np.ones(n)
            keep["radii"] = radii
        d.radii = _ptr(radii, ctypes.c_double)
        d.scaled_radii = _ptr(f64("scaled_radii"), ctypes.c_double)
        d.masses = _ptr(f64("masses"), ctypes.c_double)

        def rows(name, width):
            a = i32(name, (width,))
            return (0, None) if a is None or a.size == 0 else (a.shape[0], a)

        nb, bi = rows("bonds", 2)
        d.n_bonds, d.bond_idx = nb, _ptr(bi, ctypes.c_int32)
        d.bond_params = _ptr(f64("bond_params", (2,)), ctypes.c_double) if nb else None
        nu, ui = rows("urey_bradley_bonds", 2)
        d.n_ub, d.ub_idx = nu, _ptr(ui, ctypes.c_int32)
        d.ub_params = _ptr(f64("urey_bradley_params", (2,)), ctypes.c_double) if nu else None
        na, ai = rows("angles", 3)
        d.n_angles, d.angle_idx = na, _ptr(ai, ctypes.c_int32)
        d.angle_params = _ptr(f64("angle_params", (2,)), ctypes.c_double) if na else None
        dih_name = "dihedrals" if sysd.get("dihedrals") is not None and np.size(sysd.get("dihedrals")) else "proper_dihedrals"
        ndh, dh = rows(dih_name, 4)
        d.n_dihedrals, d.dihedral_idx = ndh, _ptr(dh, ctypes.c_int32)
        if ndh:
            dp = np.ascontiguousarray(np.asarray(sysd.get("dihedral_params", sysd.get("proper_dihedral_params")),
                                                 np.float64))
            if dp.ndim != 3 or dp.shape[-1] != 3:
                raise ValueError("dihedral_params must be (N, T, 3) [periodicity, phase, k] (bonded.py:128-137)")
            keep["dihedral_params"] = dp
            d.dihedral_terms = dp.shape[1]
            d.dihedral_params = _ptr(dp, ctypes.c_double)
        imp_name = "impropers" if sysd.get("impropers") is not None and np.size(sysd.get("impropers")) else "improper_dihedrals"
        nim, im = rows(imp_name, 4)
        d.n_impropers, d.improper_idx = nim, _ptr(im, ctypes.c_int32)
        if nim:
            ip = np.ascontiguousarray(np.asarray(sysd.get("improper_params", sysd.get("improper_dihedral_params")),
                                                 np.float64))
            if ip.ndim == 2:
                ip = ip[:, None, :]
            if ip.shape[-1] not in (2, 3):
                raise ValueError("improper_params last dimension must be 2 (harmonic) or 3 (periodic)")
            keep["improper_params"] = ip
            d.improper_terms = ip.shape[1]
            d.improper_param_width = ip.shape[-1]
            d.improper_params = _ptr(ip, ctypes.c_double)
        nc, ct = rows("cmap_torsions", 5)
        if nc and sysd.get("cmap_energy_grids") is not None:
            grids = f64("cmap_energy_grids")
            grids = grids.reshape(grids.shape[0] if grids.ndim == 3 else 1, grids.shape[-2], grids.shape[-1]) \
                if grids.ndim >= 2 else grids
            keep["cmap_energy_grids"] = grids
            d.n_cmap = nc
            d.cmap_torsions = _ptr(ct, ctypes.c_int32)
            d.cmap_map = _ptr(i32("cmap_indices"), ctypes.c_int32)
            d.n_cmap_maps = grids.shape[0]
            d.cmap_grid = grids.shape[1]
            d.cmap_grids = _ptr(grids, ctypes.c_double)
        ei = i32("excl_indices")
        if ei is not None and ei.size:
            ei = ei.reshape(n, -1)
            keep["excl_indices"] = ei
            sv = f64("excl_scales_vdw")
            se = f64("excl_scales_elec")
            sv = np.zeros(ei.shape) if sv is None else sv.reshape(ei.shape)
            se = np.zeros(ei.shape) if se is None else se.reshape(ei.shape)
            keep["excl_scales_vdw"], keep["excl_scales_elec"] = sv, se
            d.excl_cols = ei.shape[1]
            d.excl_idx = _ptr(ei, ctypes.c_int32)
            d.excl_scale_vdw = _ptr(sv, ctypes.c_double)
            d.excl_scale_elec = _ptr(se, ctypes.c_double)
        ne, ep = rows("exception_pairs", 2)
        if ne:
            d.n_exceptions = ne
            d.exc_pairs = _ptr(ep, ctypes.c_int32)
            d.exc_sigma = _ptr(f64("exception_sigmas"), ctypes.c_double)
            d.exc_epsilon = _ptr(f64("exception_epsilons"), ctypes.c_double)
            d.exc_chargeprod = _ptr(f64("exception_chargeprods"), ctypes.c_double)
        d.cutoff = float(cutoff)
        d.coulomb_constant = float(coulomb_constant)
        d.gb_coulomb_constant = float(gb_coulomb_constant)
        d.solute_dielectric = float(solute_dielectric)
        d.solvent_dielectric = float(solvent_dielectric)
        d.dielectric_offset = float(dielectric_offset)
        d.implicit_solvent = 1 if implicit_solvent else 0
        d.gb_cutoff = 1 if gb_cutoff else 0
        self.desc = d
        self.keep = keep
        self.n_atoms = n


def cmap_coefficients(grid_psi_phi) -> np.ndarray:
    """Host-side bicubic patch coefficients (G, G, 16) of one [psi, phi] grid."""
    g = np.ascontiguousarray(grid_psi_phi, dtype=np.float64)
    G = g.shape[0]
    out = np.zeros((G, G, 16))
    check(lib().prx_cmap_coefficients(_ptr(g, ctypes.c_double), G, _ptr(out, ctypes.c_double)),
          "prx_cmap_coefficients")
    return out


def plan_tiles(desc: Desc):
    ni, nx = ctypes.c_int64(), ctypes.c_int64()
    check(lib().prx_plan_tiles(ctypes.byref(desc.desc), ctypes.byref(ni), ctypes.byref(nx)), "prx_plan_tiles")
    return ni.value, nx.value
