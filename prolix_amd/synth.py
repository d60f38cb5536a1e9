"""Seeded synthetic protein-like systems for the GB-OBC2 implicit-solvent workloads.

Every system is a linear chain of residues. A residue holds the backbone atoms
N, H, CA, HA, C, O plus a linear side chain of heavy atoms carrying hydrogens.
The C-alpha trace is a self-avoiding random walk with 3.8 A steps; with
``radius`` given, the walk is confined to a sphere of that radius, which packs
the chain to protein density. Every other atom is placed 1.0-1.5 A from its
parent at a tetrahedral/trigonal angle, choosing among random candidates the
one with the most clearance from already placed atoms.

Topology follows the bonded graph: bonds (parent-child), angles (all
neighbour pairs of every atom), proper torsions (all bond paths of length 3),
one periodic improper per sp2 centre (peptide C and N), and CMAP on every
interior residue's phi/psi pair with rows [C(r-1), N(r), CA(r), C(r), N(r+1)].

Parameter distributions:
- charges N(0, 0.3) e, shifted to net zero;
- LJ sigma U(1.0, 3.5) A and epsilon U(0.01, 0.2) kcal/mol;
- OBC2 intrinsic radii U(1.2, 2.0) A and scale factors U(0.72, 0.85), so that
  ``scaled_radii = radii * scale`` is an absolute length;
- CMAP grids U(-2, 2) kcal/mol on 24x24 nodes, stored [psi, phi] as prolix
  stores them.

Bond and angle rest values equal the generated geometry, so the bonded terms
start at their minima.

Masses come from the element table that prolix uses (``constants.py:5-13``).
Units: A, kcal/mol, e, amu.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# Element masses in amu, the values of prolix's ELEMENT_MASS table (constants.py:5-13).
ELEMENT_MASS = {"H": 1.008, "C": 12.011, "N": 14.007, "O": 15.999, "S": 32.06}

BACKBONE_ATOMS_PER_RESIDUE = 6  # N, H, CA, HA, C, O
CMAP_GRID = 24


@dataclass
class SyntheticProtein:
    """Arrays in prolix's PhysicsSystem layout (typing.py:248-310), as numpy."""

    positions: np.ndarray  # (N, 3) f64
    charges: np.ndarray  # (N,)
    sigmas: np.ndarray
    epsilons: np.ndarray
    radii: np.ndarray
    scaled_radii: np.ndarray
    masses: np.ndarray
    elements: list
    bonds: np.ndarray  # (Nb, 2) int32
    bond_params: np.ndarray  # (Nb, 2) [r0, k]
    angles: np.ndarray  # (Na, 3)
    angle_params: np.ndarray  # (Na, 2) [theta0, k]
    dihedrals: np.ndarray  # (Nt, 4)
    dihedral_params: np.ndarray  # (Nt, T, 3) [n, phase, k]
    impropers: np.ndarray  # (Ni, 4)
    improper_params: np.ndarray  # (Ni, T, 3)
    cmap_torsions: np.ndarray  # (Nc, 5)
    cmap_indices: np.ndarray  # (Nc,)
    cmap_energy_grids: np.ndarray  # (M, G, G) stored [psi, phi]
    n_residues: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def n_atoms(self) -> int:
        return int(self.charges.shape[0])

    def system_dict(self, exclusions: bool = True) -> dict:
        """PhysicsSystem-style dict; with exclusions, the (N, 32) rows of the
        bond graph's 1-2/1-3 (scale 0) and 1-4 (0.5 / 0.8333) pairs."""
        d = self.as_dict()
        if exclusions:
            if "excl" not in self.meta:
                from prolix_amd.physics.neighbor_list import ExclusionSpec, map_exclusions_to_dense_padded

                spec = ExclusionSpec.from_bonds(self.bonds, self.n_atoms)
                self.meta["excl"] = map_exclusions_to_dense_padded(spec)
            d["excl_indices"], d["excl_scales_vdw"], d["excl_scales_elec"] = self.meta["excl"]
        return d

    def as_dict(self) -> dict:
        return {
            "positions": self.positions,
            "charges": self.charges,
            "sigmas": self.sigmas,
            "epsilons": self.epsilons,
            "radii": self.radii,
            "scaled_radii": self.scaled_radii,
            "masses": self.masses,
            "bonds": self.bonds,
            "bond_params": self.bond_params,
            "angles": self.angles,
            "angle_params": self.angle_params,
            "dihedrals": self.dihedrals,
            "dihedral_params": self.dihedral_params,
            "impropers": self.impropers,
            "improper_params": self.improper_params,
            "cmap_torsions": self.cmap_torsions,
            "cmap_indices": self.cmap_indices,
            "cmap_energy_grids": self.cmap_energy_grids,
        }


def _unit(v):
    n = np.linalg.norm(v)
    return v / n if n > 0 else np.array([1.0, 0.0, 0.0])


def _perp_basis(u, rng):
    a = rng.normal(size=3)
    a -= a.dot(u) * u
    v = _unit(a)
    w = np.cross(u, v)
    return v, w


class _Grid:
    """Uniform hash grid over placed atoms, for clearance queries."""

    def __init__(self, cell=4.0):
        self.cell = cell
        self.cells: dict = {}

    def _key(self, p):
        c = self.cell
        return (int(math.floor(p[0] / c)), int(math.floor(p[1] / c)), int(math.floor(p[2] / c)))

    def add(self, idx, p):
        self.cells.setdefault(self._key(p), []).append(idx)

    def near(self, p):
        kx, ky, kz = self._key(p)
        out = []
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    lst = self.cells.get((kx + dx, ky + dy, kz + dz))
                    if lst:
                        out.extend(lst)
        return out


def _side_chain_layout(c: int, rng):
    """Split c side-chain atoms into heavy atoms and hydrogens per heavy atom."""
    if c <= 0:
        return []
    heavy = max(1, int(round(c / 3.0)))
    while True:
        hyd = c - heavy
        cap = 2 * (heavy - 1) + 3
        if hyd <= cap:
            break
        heavy += 1
    hyd = c - heavy
    counts = [0] * heavy
    order = list(range(heavy - 1, -1, -1))
    k = 0
    while hyd > 0:
        j = order[k % heavy]
        lim = 3 if j == heavy - 1 else 2
        if counts[j] < lim:
            counts[j] += 1
            hyd -= 1
        k += 1
    return counts


def _build_topology(n_atoms: int, rng):
    n_res = max(3, int(round(n_atoms / 13.8)))
    while BACKBONE_ATOMS_PER_RESIDUE * n_res > n_atoms and n_res > 1:
        n_res -= 1
    sc_total = n_atoms - BACKBONE_ATOMS_PER_RESIDUE * n_res
    sc = rng.multinomial(sc_total, np.full(n_res, 1.0 / n_res)) if sc_total > 0 else np.zeros(n_res, int)
    cap = 18
    over = np.maximum(sc - cap, 0).sum()
    sc = np.minimum(sc, cap)
    while over > 0:
        j = int(rng.integers(n_res))
        if sc[j] < cap:
            sc[j] += 1
            over -= 1
    elements: list = []
    parent: list = []
    role: list = []  # (residue, name)
    bb = []  # per residue dict name->index
    for r in range(n_res):
        d = {}

        def add(el, par, name):
            elements.append(el)
            parent.append(par)
            role.append((r, name))
            return len(elements) - 1

        d["N"] = add("N", bb[r - 1]["C"] if r > 0 else -1, "N")
        d["H"] = add("H", d["N"], "H")
        d["CA"] = add("C", d["N"], "CA")
        d["HA"] = add("H", d["CA"], "HA")
        layout = _side_chain_layout(int(sc[r]), rng)
        prev = d["CA"]
        for hi, nh in enumerate(layout):
            last = hi == len(layout) - 1
            el = "C"
            if last and len(layout) > 1:
                u = rng.random()
                el = "O" if u < 0.2 else ("N" if u < 0.35 else ("S" if u < 0.4 else "C"))
            h = add(el, prev, "SC")
            for _ in range(nh):
                add("H", h, "SCH")
            prev = h
        d["C"] = add("C", d["CA"], "C")
        d["O"] = add("O", d["C"], "O")
        bb.append(d)
    assert len(elements) == n_atoms, (len(elements), n_atoms)
    return n_res, elements, np.asarray(parent), role, bb


def _ca_walk(n_res, rng, radius):
    ca = np.zeros((n_res, 3))
    if radius is not None:
        ca[0] = rng.normal(size=3) * radius * 0.3
    prev_dir = _unit(rng.normal(size=3))
    for i in range(1, n_res):
        best, best_score = None, -1e30
        for _ in range(200):
            d = _unit(rng.normal(size=3))
            if i > 1:
                c = -prev_dir.dot(d)  # cos of the CA(i-1)-CA(i)-CA(i+1) angle
                if not (-0.87 < c < 0.09):  # angle in (85, 150) degrees
                    continue
            p = ca[i - 1] + 3.8 * d
            score = 0.0
            if radius is not None:
                rr = np.linalg.norm(p)
                if rr > radius:
                    score -= (rr - radius) * 10.0
            if i > 1:
                dmin = np.min(np.linalg.norm(ca[: i - 1] - p, axis=1))
                score += min(dmin - 4.2, 0.0) * 5.0 + min(dmin, 6.0) * 0.01
            if score > best_score:
                best, best_score = d, score
            if score >= 0.0 and radius is None:
                break
            if score > 0.0 and i > 1 and radius is not None and rng.random() < 0.5:
                break
        if best is None:
            best = _unit(rng.normal(size=3))
        ca[i] = ca[i - 1] + 3.8 * best
        prev_dir = best
    return ca


# Equilibrium bond angles are kept in [80, 150] degrees: a torsion's gradient
# diverges as either of its bond angles reaches 180 degrees, and an equilibrium
# angle near 180 (which free placement sometimes produces) puts thermal
# motion on that singularity.
ANGLE_MIN_DEG, ANGLE_MAX_DEG = 80.0, 150.0


def _place_child(pos, grid, placed_mask, p_idx, g_pos, bond, angle_deg, sig, idx, skip, rng, ntry=24,
                 close13=()):
    """Place atom ``idx`` bonded to ``p_idx`` at ``bond`` A, angle to g_pos ~ angle_deg."""
    p = pos[p_idx]
    if g_pos is None:
        dirs = rng.normal(size=(ntry, 3))
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    else:
        u = _unit(g_pos - p)
        v, w = _perp_basis(u, rng)
        th = np.deg2rad(angle_deg + rng.uniform(-8.0, 8.0, size=ntry))
        ph = rng.uniform(0.0, 2.0 * math.pi, size=ntry)
        dirs = (np.cos(th)[:, None] * u[None, :]
                + (np.sin(th) * np.cos(ph))[:, None] * v[None, :]
                + (np.sin(th) * np.sin(ph))[:, None] * w[None, :])
    cand = p[None, :] + bond * dirs
    near = [j for j in grid.near(p) if j not in skip]
    if not near:
        return cand[0]
    close = set(close13)
    is13 = np.array([j in close for j in near])
    near = np.asarray(near)
    d = np.linalg.norm(cand[:, None, :] - pos[near][None, :, :], axis=2)  # (ntry, k)
    need = np.maximum(2.2, 0.85 * 0.5 * (sig[idx] + sig[near]))
    need = np.where(is13, 1.6, need)
    score = np.min(d / need[None, :], axis=1)
    best = int(np.argmax(score))
    if score[best] < 1.0 and g_pos is not None and ntry < 96:
        # Crowded cone: widen the angle jitter and try again.
        alt = _place_child(pos, grid, placed_mask, p_idx, None, bond, angle_deg, sig, idx, skip, rng,
                           ntry=96, close13=close13)
        da = np.linalg.norm(alt[None, :] - pos[near], axis=1)
        if np.min(da / need) > score[best]:
            return alt
    return cand[best]


def make_protein(n_atoms: int, seed: int, radius: float | None = None) -> SyntheticProtein:
    """Build one seeded synthetic protein of exactly ``n_atoms`` atoms."""
    if n_atoms < 6:
        raise ValueError("make_protein needs at least 6 atoms (one backbone residue)")
    rng = np.random.default_rng(seed)
    n_res, elements, parent, role, bb = _build_topology(n_atoms, rng)
    N = n_atoms

    q = rng.normal(0.0, 0.3, size=N)
    q -= q.mean()
    sig = rng.uniform(1.0, 3.5, size=N)
    eps = rng.uniform(0.01, 0.2, size=N)
    radii = rng.uniform(1.2, 2.0, size=N)
    scale = rng.uniform(0.72, 0.85, size=N)
    masses = np.array([ELEMENT_MASS[e] for e in elements])

    children = [[] for _ in range(N)]
    for i, p in enumerate(parent):
        if p >= 0:
            children[p].append(i)

    adj0 = [[] for _ in range(N)]
    for i, p in enumerate(parent):
        if p >= 0:
            adj0[p].append(i)
            adj0[i].append(int(p))

    pos = np.full((N, 3), np.nan)
    grid = _Grid()
    placed = np.zeros(N, bool)

    def put(i, p):
        pos[i] = p
        placed[i] = True
        grid.add(i, p)

    ca = _ca_walk(n_res, rng, radius)
    for r in range(n_res):
        put(bb[r]["CA"], ca[r])
    # Backbone C(r) and N(r+1) between consecutive C-alphas.
    b_cac, b_cn, b_nca = 1.52, 1.33, 1.46
    for r in range(n_res - 1):
        a, b = ca[r], ca[r + 1]
        u = _unit(b - a)
        dab = np.linalg.norm(b - a)
        d_cb = math.sqrt(b_cn**2 + b_nca**2 - 2 * b_cn * b_nca * math.cos(math.radians(122.0)))
        cosang = (b_cac**2 + dab**2 - d_cb**2) / (2 * b_cac * dab)
        cosang = max(-1.0, min(1.0, cosang))
        alpha = math.acos(cosang)
        v, w = _perp_basis(u, rng)
        ph = rng.uniform(0, 2 * math.pi) + np.linspace(0.0, 2 * math.pi, 36, endpoint=False)
        dirs_c = (math.cos(alpha) * u[None, :] + math.sin(alpha) * (np.cos(ph)[:, None] * v[None, :]
                                                                      + np.sin(ph)[:, None] * w[None, :]))
        if placed[bb[r]["N"]]:
            # choose the phase that puts the N-CA-C angle nearest 111 degrees
            nv = _unit(pos[bb[r]["N"]] - a)
            dirc = dirs_c[int(np.argmin(np.abs(dirs_c @ nv - math.cos(math.radians(111.0)))))]
        else:
            dirc = dirs_c[0]
        cpos = a + b_cac * dirc
        put(bb[r]["C"], cpos)
        ax = b - cpos
        dd = np.linalg.norm(ax)
        ax = ax / dd
        x = (dd**2 + b_cn**2 - b_nca**2) / (2 * dd)
        h = math.sqrt(max(b_cn**2 - x**2, 0.0))
        v2, w2 = _perp_basis(ax, rng)
        ph2 = rng.uniform(0, 2 * math.pi) + np.linspace(0.0, 2 * math.pi, 36, endpoint=False)
        n_cands = cpos[None, :] + x * ax[None, :] + h * (np.cos(ph2)[:, None] * v2[None, :]
                                                       + np.sin(ph2)[:, None] * w2[None, :])
        # the phase that puts the CA-C-N angle nearest 116 degrees
        uc = _unit(a - cpos)
        un = (n_cands - cpos[None, :]) / np.linalg.norm(n_cands - cpos[None, :], axis=1, keepdims=True)
        npos = n_cands[int(np.argmin(np.abs(un @ uc - math.cos(math.radians(116.0)))))]
        put(bb[r + 1]["N"], npos)
    # Chain termini.
    first_n = bb[0]["N"]
    put(first_n, _place_child(pos, grid, placed, bb[0]["CA"], ca[1] if n_res > 1 else None,
                              b_nca, 110.0, sig, first_n, {bb[0]["CA"]}, rng))
    last_c = bb[-1]["C"]
    put(last_c, _place_child(pos, grid, placed, bb[-1]["CA"], pos[bb[-1]["N"]], b_cac, 110.0, sig,
                             last_c, {bb[-1]["CA"]}, rng))

    # Remaining atoms in index order (parents always precede children except the
    # backbone atoms placed above).
    for i in range(N):
        if placed[i]:
            continue
        p = parent[i]
        el = elements[i]
        gp = parent[p] if p >= 0 else -1
        if gp < 0 or not placed[gp]:
            # choose any placed neighbour of p other than i as the angle reference
            refs = [c for c in children[p] if placed[c] and c != i]
            gpos = pos[refs[0]] if refs else None
        else:
            gpos = pos[gp]
        name = role[i][1]
        if el == "H":
            bond = rng.uniform(1.0, 1.1)
        else:
            bond = rng.uniform(1.3, 1.5)
        if name == "O":
            bond = 1.23
        ang = 120.0 if name in ("O", "H") else 109.5
        skip = {p}
        close13 = [c for c in children[p] if c != i]
        if gp >= 0:
            close13.append(gp)
        elif p >= 0 and parent[p] < 0:
            pass
        put(i, _place_child(pos, grid, placed, p, gpos, bond, ang, sig, i, skip, rng,
                            close13=close13 + [c for c in adj0[p] if c != i]))

    # ---- topology lists ----
    bonds = np.array([(parent[i], i) for i in range(N) if parent[i] >= 0], dtype=np.int32)
    adj = [[] for _ in range(N)]
    for a, b in bonds:
        adj[a].append(int(b))
        adj[b].append(int(a))
    angles = []
    for j in range(N):
        nb = sorted(adj[j])
        for x in range(len(nb)):
            for y in range(x + 1, len(nb)):
                angles.append((nb[x], j, nb[y]))
    angles = np.array(angles, dtype=np.int32).reshape(-1, 3)
    tors = []
    for j, k in bonds:
        for i in adj[j]:
            if i == k:
                continue
            for l in adj[k]:
                if l == j or l == i:
                    continue
                tors.append((i, j, k, l))
    tors = np.array(tors, dtype=np.int32).reshape(-1, 4)
    imps = []
    for r in range(n_res):
        c = bb[r]["C"]
        if r + 1 < n_res:
            imps.append((bb[r]["CA"], bb[r + 1]["N"], c, bb[r]["O"]))
        if r > 0:
            imps.append((bb[r - 1]["C"], bb[r]["CA"], bb[r]["N"], bb[r]["H"]))
    imps = np.array(imps, dtype=np.int32).reshape(-1, 4)
    cmap_rows = []
    for r in range(1, n_res - 1):
        cmap_rows.append((bb[r - 1]["C"], bb[r]["N"], bb[r]["CA"], bb[r]["C"], bb[r + 1]["N"]))
    cmap_rows = np.array(cmap_rows, dtype=np.int32).reshape(-1, 5)

    # ---- parameters ----
    d_b = np.linalg.norm(pos[bonds[:, 0]] - pos[bonds[:, 1]], axis=1)
    bond_params = np.stack([d_b, rng.uniform(500.0, 800.0, size=len(bonds))], axis=1)
    if len(angles):
        v1 = pos[angles[:, 0]] - pos[angles[:, 1]]
        v2 = pos[angles[:, 2]] - pos[angles[:, 1]]
        ct = np.sum(v1 * v2, 1) / (np.linalg.norm(v1, axis=1) * np.linalg.norm(v2, axis=1))
        th0 = np.clip(np.arccos(np.clip(ct, -1.0, 1.0)), math.radians(ANGLE_MIN_DEG), math.radians(ANGLE_MAX_DEG))
    else:
        th0 = np.zeros(0)
    angle_params = np.stack([th0, rng.uniform(80.0, 140.0, size=len(angles))], axis=1)
    T = 2
    nt = len(tors)
    dihedral_params = np.zeros((nt, T, 3))
    dihedral_params[:, :, 0] = rng.integers(1, 4, size=(nt, T))
    dihedral_params[:, :, 1] = rng.choice([0.0, math.pi], size=(nt, T))
    dihedral_params[:, :, 2] = rng.uniform(0.0, 0.8, size=(nt, T))
    ni = len(imps)
    improper_params = np.zeros((ni, 1, 3))
    improper_params[:, 0, 0] = 2.0
    improper_params[:, 0, 1] = math.pi
    improper_params[:, 0, 2] = rng.uniform(1.0, 10.0, size=ni)
    n_maps = 4
    grids = rng.uniform(-2.0, 2.0, size=(n_maps, CMAP_GRID, CMAP_GRID))
    cmap_idx = rng.integers(0, n_maps, size=len(cmap_rows)).astype(np.int32)

    return SyntheticProtein(
        positions=pos,
        charges=q,
        sigmas=sig,
        epsilons=eps,
        radii=radii,
        scaled_radii=radii * scale,
        masses=masses,
        elements=list(elements),
        bonds=bonds,
        bond_params=bond_params,
        angles=angles,
        angle_params=angle_params,
        dihedrals=tors,
        dihedral_params=dihedral_params,
        impropers=imps,
        improper_params=improper_params,
        cmap_torsions=cmap_rows,
        cmap_indices=cmap_idx,
        cmap_energy_grids=grids,
        n_residues=n_res,
        meta={"seed": seed, "radius": radius},
    )


def compact_radius(n_atoms: int) -> float:
    """Sphere radius that packs n atoms like the 2,500-atom, 2.0 nm globule."""
    return 20.0 * (n_atoms / 2500.0) ** (1.0 / 3.0)


def peptide_138(seed: int = 0) -> SyntheticProtein:
    """The CPU-runnable 10-residue, 138-atom peptide (unconfined random walk)."""
    return make_protein(138, seed, radius=None)


def globule_2500(seed: int = 1) -> SyntheticProtein:
    """The 2,500-atom globular protein packed into a 20 A sphere."""
    return make_protein(2500, seed, radius=compact_radius(2500))


def batch_mix_sizes(n_systems: int = 256, lo: int = 300, hi: int = 5000, seed: int = 2) -> np.ndarray:
    """Atom counts drawn log-uniform in [lo, hi] (seeded)."""
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size=n_systems)).astype(np.int64)


def _mix_member(args):
    n, sd = args
    p = make_protein(int(n), sd, radius=compact_radius(int(n)))
    p.system_dict(exclusions=True)  # build the exclusion rows in the worker
    return p


def build_many(jobs, workers: int | None = None, start_method: str = "fork"):
    """Compact proteins for (n_atoms, seed) jobs, built in a process pool
    (``workers`` processes, or the CPUs this process may use, at most 16). With
    the default "fork" start method call it before the process touches the GPU;
    "spawn" starts fresh interpreters and is safe afterwards. Each system
    depends only on its own seed, so the result is the same for any worker
    count."""
    import os

    jobs = [(int(n), int(sd)) for n, sd in jobs]
    if workers is None:
        workers = min(16, len(os.sched_getaffinity(0)))
    if workers <= 1 or len(jobs) <= 1:
        return [_mix_member(j) for j in jobs]
    import multiprocessing as mp

    # largest first, so the pool's tail is short
    order = sorted(range(len(jobs)), key=lambda k: -jobs[k][0])
    with mp.get_context(start_method).Pool(min(workers, len(jobs))) as pool:
        built = pool.map(_mix_member, [jobs[k] for k in order], chunksize=1)
    systems = [None] * len(jobs)
    for k, p in zip(order, built):
        systems[k] = p
    return systems


def mix_jobs(n_systems: int = 256, lo: int = 300, hi: int = 5000, seed: int = 2):
    """(n_atoms, seed) jobs of the uneven many-system mix."""
    sizes = batch_mix_sizes(n_systems, lo, hi, seed)
    return [(int(n), seed * 100003 + k) for k, n in enumerate(sizes)]


def mix_temperatures(n_systems: int = 256, seed: int = 2, temperatures=(280.0, 350.0)) -> np.ndarray:
    rng = np.random.default_rng(seed + 7919)
    return rng.uniform(temperatures[0], temperatures[1], size=n_systems)


def batch_mix(n_systems: int = 256, lo: int = 300, hi: int = 5000, seed: int = 2,
              temperatures=(280.0, 350.0), workers: int | None = None):
    """The uneven many-system mix: proteins plus per-system temperatures."""
    systems = build_many(mix_jobs(n_systems, lo, hi, seed), workers)
    return systems, mix_temperatures(n_systems, seed, temperatures)
