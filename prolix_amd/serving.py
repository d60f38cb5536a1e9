"""Continuous-batching request loop over a GBPool.

Requests (a system, a step count, a temperature, a Philox seed) arrive over
time. The loop admits every arrived request whose estimated device footprint
fits the HBM budget (and the pool's arenas), advances all live systems
together in chunks of steps (one captured graph per chunk), and as soon as a
system has taken its steps reads its state back and evicts it, so waiting
requests take its place while the others keep running. Positions, momenta,
forces, Born radii and Philox counters stay on the device for a system's whole
life. Reported: per-request completion latency (finish - arrival), p50/p99,
aggregate ns/day and pair interactions/s.

prolix has no serving loop; its batched path pads systems into buckets and
runs them to completion together (padding.py:573-654,
batched_simulate.py:1125-1422).
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from prolix_amd.engine import GBPool, Langevin

TILE = 64


def hbm_bytes(n_atoms: int, n_bonded: int = 0) -> int:
    """Device footprint of one system: per-atom arrays (~330 B), bonded entries
    with force records and energy slots (~120 B each) and the pair work list
    with its 2 x 64 float4 partial slots and energies (~2.1 KB per 64x64 tile pair)."""
    nt = (n_atoms + TILE - 1) // TILE
    return 330 * n_atoms + 120 * n_bonded + 2100 * nt * (nt + 1) // 2


@dataclass
class Request:
    system: dict
    n_steps: int
    temperature: float
    seed: int
    arrival: float
    dt_fs: float = 1.0
    gamma_ps: float = 1.0
    rid: int = 0


@dataclass
class Result:
    rid: int
    n_atoms: int
    n_steps: int
    arrival: float
    admitted: float
    finished: float
    positions: np.ndarray
    momenta: np.ndarray
    forces: np.ndarray
    energies: np.ndarray | None = None
    extra: dict = field(default_factory=dict)
    error: str | None = None  # set when the system was evicted early (non-finite state)

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def _bonded(sysd) -> int:
    n = 0
    for k in ("bonds", "angles", "dihedrals", "impropers", "cmap_torsions", "urey_bradley_bonds"):
        v = sysd.get(k)
        if v is not None:
            n += len(v) * (2 if k == "dihedrals" else 1)
    return n


def serve(requests, hbm_budget_bytes: float = 200e9, max_systems: int = 1024, chunk_steps: int = 100,
          max_atoms: int | None = None, sample_energies=(), safety_caps: bool = True, clock=None, log=None,
          chunk_s: float | None = None):
    """Run the request trace to completion; returns (results, stats).

    Each admitted system carries its step count as a device step limit
    (GBPool.set_step_limit): a chunk may run past the end of a request without
    changing its result, so chunks need not stop at the next request's end.
    Admission queues its uploads behind the running chunk without waiting for
    it, so while a chunk runs the loop admits as many arrived requests as
    systems finish in that chunk: they join the next chunk, which they start
    with no more than max_systems systems stepping. After each chunk the
    finished (and failed) systems are read back and evicted and further
    arrived requests admitted; the pool computes the new systems' momenta and
    forces together at the start of the next chunk.

    chunk_steps: steps per chunk, or with chunk_s the longest chunk; chunk_s:
    the target wall time of a chunk, whose length then follows the step rate
    of the chunk before it (in multiples of 50 steps, the engine's graph).
    """
    import torch

    reqs = sorted(requests, key=lambda r: r.arrival)
    if any(r.n_steps < 1 for r in reqs):
        raise ValueError("every request needs n_steps >= 1")
    if max_atoms is None:
        max_atoms = max(1, int(min(sum(len(r.system["charges"]) for r in reqs), 8_000_000)))
    # device memory, measured: the pool's arenas at creation, then the peak
    # over the run (the work lists and partial sums grow with the live set)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    pool = GBPool(max_atoms=max_atoms, max_systems=max_systems)
    torch.cuda.synchronize()
    pool_bytes = free0 - torch.cuda.mem_get_info()[0]
    peak_bytes = pool_bytes
    clock = clock or time.perf_counter
    t0 = clock()
    now = lambda: clock() - t0
    waiting = list(reqs)
    live = {}  # slot -> [request, remaining, admitted time, footprint]
    results = []
    used = 0
    peak_used = 0
    pair_steps = 0.0
    sys_steps = 0
    chunks = 0
    chunk = chunk_steps
    steady = None  # (wall, system-steps, pair-steps) when the last request was admitted
    t_admit = t_launch = t_wait = t_evict = 0.0  # host wall time spent in each phase
    sample = set(sample_energies)
    descs = {}  # rid -> host-side description, made while the device steps
    def admit_arrived(limit):
        """Admit arrived requests in arrival order, at most limit, while the
        budget and the pool allow; returns the slots admitted."""
        nonlocal used, peak_used
        got = []
        t = now()
        while waiting and waiting[0].arrival <= t and len(got) < limit:
            r = waiting[0]
            cost = hbm_bytes(len(r.system["charges"]), _bonded(r.system))
            if used + cost > hbm_budget_bytes and live:
                break
            slot = pool.admit(r.system, Langevin(r.dt_fs, r.temperature, r.gamma_ps, r.seed, safety_caps),
                              desc=descs.pop(r.rid, None))
            if slot < 0:
                if not live:
                    raise RuntimeError(f"request {r.rid} ({len(r.system['charges'])} atoms) does not fit the pool")
                break
            pool.set_step_limit(slot, r.n_steps)
            waiting.pop(0)
            used += cost
            peak_used = max(peak_used, used)
            live[slot] = [r, r.n_steps, now(), cost]
            got.append(slot)
        return got

    while waiting or live:
        ta = clock()
        admit_arrived(max_systems - len(live))
        t_admit += clock() - ta
        if steady is None and not waiting:
            steady = (now(), sys_steps, pair_steps)
        if not live:
            wait = waiting[0].arrival - now()
            if wait > 0:
                time.sleep(min(wait, 0.05))
            continue
        ts = clock()
        pool.steps(chunk)
        tl = clock()
        # while the chunk runs: admit as many arrived requests as systems
        # finish in it (they join the next chunk), then describe the requests
        # likely admitted after it
        stepping = set(live)
        n_fin = sum(1 for v in live.values() if v[1] <= chunk)
        admit_arrived(n_fin)
        for r in waiting[:n_fin + 1]:
            if r.rid not in descs and r.arrival <= now() + 0.5:
                descs[r.rid] = GBPool.describe(r.system)
        tj = clock()
        torch.cuda.synchronize()
        tw = clock()
        peak_bytes = max(peak_bytes, free0 - torch.cuda.mem_get_info()[0])
        t_launch += tl - ts
        t_admit += tj - tl
        t_wait += tw - tj
        chunks += 1
        te = clock()
        for slot, v in live.items():
            if slot not in stepping:
                continue  # admitted during the chunk
            k = min(chunk, v[1])  # steps the system took (it stops at its limit)
            v[1] -= k
            n = len(v[0].system["charges"])
            pair_steps += k * n * (n - 1) / 2.0
            sys_steps += k
        # explosion check on the device counters: a system whose positions or
        # momenta turned non-finite is evicted with an error result; the others
        # never read its atoms, so they continue unchanged
        failed = {slot for slot, c in pool.nonfinite_counts().items() if c > 0}
        for slot in [s for s, v in live.items() if s in stepping and (v[1] <= 0 or s in failed)]:
            r, _, adm, cost = live.pop(slot)
            x, p, f, steps = pool.read(slot)
            err = None
            if slot in failed:
                err = f"non-finite positions or momenta after {steps} of {r.n_steps} steps"
            e = pool.energies(slot) if r.rid in sample and err is None else None
            results.append(Result(r.rid, len(r.system["charges"]), steps, r.arrival, adm, now(), x, p, f, e,
                                  error=err))
            pool.release(slot)
            used -= cost
            if log:
                log(f"request {r.rid}: {len(r.system['charges'])} atoms, {steps} steps, "
                    f"latency {results[-1].latency:.2f}s, live {len(live)}, waiting {len(waiting)}"
                    + (f", evicted: {err}" if err else ""))
        t_evict += clock() - te
        if chunk_s is not None:
            rate = chunk / max(tw - ts, 1e-6)  # steps per second of the last chunk
            chunk = int(min(chunk_steps, max(50, 50 * round(rate * chunk_s / 50.0))))
    wall = now()
    pool.close()
    lat = np.array([r.latency for r in results])
    st_wall, st_sys, st_pair = steady if steady is not None else (wall, sys_steps, pair_steps)
    stats = {
        "requests": len(results),
        "failed": sum(r.error is not None for r in results),
        "wall_s": wall,
        "chunks": chunks,
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "aggregate_ns_per_day": sys_steps * 1e-6 * 86400.0 / wall * (reqs[0].dt_fs if reqs else 1.0),
        "pair_interactions_per_s": pair_steps / wall,
        "offered_pair_steps": float(sum(len(r.system["charges"]) ** 2 / 2.0 * r.n_steps for r in reqs)),
        # up to the admission of the last request: while requests were still
        # waiting, so the pool held as many systems as the limits allow
        "until_last_admission": {"wall_s": st_wall,
                                 "aggregate_ns_per_day": st_sys * 1e-6 * 86400.0 / max(st_wall, 1e-9)
                                 * (reqs[0].dt_fs if reqs else 1.0),
                                 "pair_interactions_per_s": st_pair / max(st_wall, 1e-9)},
        # admit: host work of admissions (those made while a chunk ran overlap
        # it); launch: momenta, forces, work lists and graphs of the admitted
        # systems plus the enqueue of the chunk; wait: until the chunk ends;
        # evict: counters, read-back and release
        "host_time_s": {"admit": t_admit, "launch": t_launch, "wait": t_wait, "evict": t_evict},
        # measured device bytes: the pool's arenas (sized by max_atoms at
        # creation) and the peak with the work lists of the live systems; the
        # budget admits requests by their estimated work-list footprint
        "device_bytes": {"budget": float(hbm_budget_bytes), "pool_at_creation": int(pool_bytes),
                         "peak": int(peak_bytes), "estimated_peak_admitted": int(peak_used)},
    }
    return sorted(results, key=lambda r: r.rid), stats


def poisson_trace(systems, rate_per_s: float, steps_range, temps_range=(280.0, 350.0), seed: int = 4,
                  dt_fs: float = 1.0, gamma_ps: float = 1.0):
    """Seeded finite trace: Poisson arrivals, uniform step counts and temperatures, one seed each."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_per_s, size=len(systems)))
    t -= t[0]
    steps = rng.integers(steps_range[0], steps_range[1] + 1, size=len(systems))
    temps = rng.uniform(temps_range[0], temps_range[1], size=len(systems))
    return [Request(s, int(steps[k]), float(temps[k]), int(rng.integers(1, 2**62)), float(t[k]), dt_fs, gamma_ps, k)
            for k, s in enumerate(systems)]
