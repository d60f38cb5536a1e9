"""Benchmark: batched GB-OBC2 implicit-solvent MD on one MI355X.

Workload: 256 independent synthetic proteins, atom counts log-uniform in
[300, 5000] (seeded), each a confined random-walk chain packed to protein
density; all-pairs OBC2 Born radii + GB + Coulomb + LJ (no cutoff), bonded
terms and CMAP; Langevin BAOAB at dt = 1 fs, gamma = 1/ps, per-system
temperatures U(280, 350) K, with the batched step's safety caps. One "step"
advances every system by one BAOAB step.

Printed JSON line: aggregate ns/day over all systems (value), pair
interactions/s, the roofline of the dominant pair kernel (HIP-event timed),
the plain one-system-per-engine path on the same mix, and a CPU baseline
(the C++/OpenMP restatement, float32) on a bounded sample. Beside them:
- "serving": a finite seeded Poisson trace (50 requests/s, N log-uniform in
  [200, 8000], n_steps uniform in [1000, 50000], T U(280, 350) K, own Philox
  seed each) replayed through the continuous-batching loop: p50/p99
  completion latency, aggregate ns/day, offered vs achieved pair-steps/s
  (that trace offers far more than the device achieves, so its latencies are
  queueing); "serving.sustained": the same sizes at 8 requests/s with n_steps
  U[1000, 5000], a load the device sustains, p50/p99 there;
  "serving.mix": the batched mix served (512 requests, 256 live at most),
  its ns/day against the batched figure;
- "large_cutoff": one 20,000-atom globule with a 20 A cutoff and the
  neighbour-list GB form, ms/step and interactions/s within the cutoff.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "aggregate ns/day + pair-interactions/s, batched GB-OBC2 MD; p50/p99 latency"
FLOPS_PER_PAIR = {"born_pairs": 56, "force_pairs": 73, "chain_pairs": 62}  # algorithmic, per unordered pair
FLOPS_TOTAL = 191
PEAK_FP32 = 157.3  # TFLOP/s, MI355X FP32 vector peak (packed FP32)
PEAK_FP32_UNPACKED = 78.6  # TFLOP/s, one FP32 FMA per lane per clock (the kernels issue above it: plain
#                            v_fma_f32 measured at 1.36 wave64 instructions per clock per CU)
DT_FS = 1.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def ns_per_day(steps_per_s):
    return steps_per_s * DT_FS * 1e-6 * 86400.0


def cpu_baseline(systems, sizes_full, budget_s=20.0, n_sample=16, repeats=5):
    """Time the C++ restatement (float32, all host cores) on a seeded subset of
    the same mix and scale by sum N^2 to the whole mix: one warm-up step, then
    the median of ``repeats`` timed runs sized to share ``budget_s``."""
    from oracle import cpu as C

    rng = np.random.default_rng(11)
    idx = rng.choice(len(systems), size=min(n_sample, len(systems)), replace=False)
    sub = [systems[k] for k in idx]
    cs = [C.CpuSystem(s) for s in sub]
    states = []
    for s, c in zip(sub, cs):
        x = np.ascontiguousarray(s["positions"], np.float32)
        _, f = c.energy_forces(x, np.float32)
        states.append([x, np.zeros_like(x), f])

    def run(steps):
        t = time.perf_counter()
        for c, (x, p, f) in zip(cs, states):
            c.baoab(x, p, f, steps, DT_FS, 300.0, 1.0, 7, dtype=np.float32)
        return time.perf_counter() - t

    t1 = run(1)
    steps = max(1, min(50, int(budget_s / repeats / max(t1, 1e-3))))
    t = float(np.median([run(steps) for _ in range(repeats)]))
    sub_sizes = np.array([len(s["charges"]) for s in sub], np.float64)
    sub_pairs = float(np.sum(sub_sizes * (sub_sizes - 1) / 2))
    full_pairs = float(np.sum(sizes_full * (sizes_full - 1.0) / 2))
    sec_per_full_step = t / steps * full_pairs / sub_pairs
    return {
        "value": ns_per_day(1.0 / sec_per_full_step) * len(sizes_full),
        "unit": "ns/day",
        "cores": C.threads(),
        "kind": "port",
        "sample": (f"{len(sub)} of the {len(sizes_full)} systems ({int(sub_sizes.sum())} atoms), median of "
                   f"{repeats} runs of {steps} BAOAB steps each in float32 ({t:.2f} s per run, after one warm-up "
                   "step); scaled to the full mix by sum N(N-1)/2"),
        "pair_interactions_per_s": sub_pairs * steps / t,
    }


def pass_timings(eng, n_steps):
    """Per-pass mean durations (ms) inside the step graph: a one-lane kernel
    on the pass's own stream stores the GPU wall clock before and after every
    pass (HIP events recorded inside a captured graph cannot be timed on this
    ROCm). Falls back to HIP events around back-to-back launches of each pass."""
    try:
        return eng.profile_steps(n_steps), (f"GPU wall-clock stamps around each pass inside the step graph, "
                                            f"{n_steps} integrator steps after the timed region")
    except Exception as exc:  # noqa: BLE001 - report and fall back
        log(f"in-graph pass timing unavailable ({exc}); timing back-to-back launches")
        return eng.profile_passes(10), "HIP events around 10 back-to-back launches of each pass"


def single_system(sysd, gpu_steps, label):
    """One system alone on the plain GPU path (its own engine, the default
    step path: the fused kernel for small systems, graph-captured steps
    otherwise), and each step path on request. Returns the result and the
    relaxed start positions, from which the CPU baseline leg times the C++
    restatement on the same system."""
    import torch

    from prolix_amd.engine import GBEngine, Langevin
    from prolix_amd.minimize import minimize

    n = len(sysd["charges"])
    eng = GBEngine(sysd, cutoff=0.0)
    x = minimize(eng, sysd["positions"], steps=100)

    def timed(mode, steps):
        eng.set_step_mode(mode)
        eng.init(x, Langevin(DT_FS, 300.0, 1.0, seed=3, safety_caps=True))
        eng.steps(100)
        torch.cuda.synchronize()
        t = time.perf_counter()
        eng.steps(steps)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps

    g = timed("auto", gpu_steps)  # the engine's default path
    modes = {m: timed(m, min(gpu_steps, 2000)) * 1e3 for m in ("graph", "persistent", "fused")}
    eng.close()
    out = {"workload": label, "atoms": n, "gpu_steps": gpu_steps, "gpu_ms_per_step": g * 1e3,
           "gpu_ns_per_day": ns_per_day(1.0 / g), "gpu_pair_interactions_per_s": n * (n - 1) / 2 / g,
           "ms_per_step_by_step_mode": modes}
    return out, np.ascontiguousarray(x.cpu().numpy(), np.float32)


def cpu_single(sysd, x, cpu_steps, gpu_steps, repeats=3):
    """The C++ restatement on one system from the same start (CPU baseline leg)."""
    from oracle import cpu as C

    c = C.CpuSystem(sysd)
    xc = x.copy()
    _, f = c.energy_forces(xc, np.float32)
    pc = np.zeros_like(xc)
    c.baoab(xc, pc, f, 1, DT_FS, 300.0, 1.0, 7, dtype=np.float32)  # warm-up

    def run():
        t0 = time.perf_counter()
        c.baoab(xc, pc, f, cpu_steps, DT_FS, 300.0, 1.0, 7, dtype=np.float32)
        return time.perf_counter() - t0

    tc = float(np.median([run() for _ in range(repeats)])) / cpu_steps
    return {"cpu_steps": cpu_steps, "cpu_ms_per_step": tc * 1e3, "cpu_ns_per_day": ns_per_day(1.0 / tc),
            "cpu_cores": C.threads(), "cpu_kind": "port",
            "cpu_sample": f"median of {repeats} runs of {cpu_steps} float32 BAOAB steps after one warm-up"
                          + ("" if cpu_steps >= gpu_steps else f"; per-step time extrapolated to {gpu_steps} steps")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--n-systems", type=int, default=256)
    ap.add_argument("--min-steps", type=int, default=2000,
                    help="steepest-descent steps before dynamics (mix and the cutoff system)")
    ap.add_argument("--lbfgs-steps", type=int, default=2000,
                    help="L-BFGS steps after the steepest descent (the batched minimiser's last stage)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--sustained-requests", type=int, default=120)
    ap.add_argument("--sustained-rate", type=float, default=8.0, help="requests/s of the sustained-load trace")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--no-single", action="store_true",
                    help="skip the one-system legs (138-atom peptide, 2,500-atom globule)")
    ap.add_argument("--calibration-copy-mb", type=float, default=0.0,
                    help="also run one elementwise copy of this many MiB (float32), the reference kernel "
                         "against which rocprofv3 FETCH_SIZE / WRITE_SIZE readings are scaled")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}): there is no multi-GPU path; "
              "independent systems would shard by whole systems, one engine per GPU.", file=sys.stderr)
        sys.exit(2)

    # 1. synthetic inputs, generated before the process touches the GPU (forked pool)
    from prolix_amd import synth

    t0 = time.time()
    jobs = synth.mix_jobs(a.n_systems, 300, 5000, seed=2)
    n_mix = len(jobs)
    if not a.no_serve:
        jobs += synth.mix_jobs(a.serve_requests, 200, 8000, seed=5)
        jobs += synth.mix_jobs(a.sustained_requests, 200, 8000, seed=8)
    if not a.no_large:
        jobs.append((20000, 7))
    built = synth.build_many(jobs)
    prots, temps = built[:n_mix], synth.mix_temperatures(a.n_systems, seed=2)
    serve_prots = built[n_mix:n_mix + (0 if a.no_serve else a.serve_requests)]
    sustained_prots = built[n_mix + len(serve_prots):n_mix + len(serve_prots)
                            + (0 if a.no_serve else a.sustained_requests)]
    large = None if a.no_large else built[-1]
    singles = []
    if not a.no_single:
        singles = [("peptide: 138 atoms (10 residues), no cutoff, BAOAB 300 K, gamma 1/ps, dt 1 fs",
                    synth.make_protein(138, 0).system_dict(), 1000, 1000),
                   ("globule: 2,500 atoms compacted to a 20 A radius, no cutoff, BAOAB 300 K, gamma 1/ps, dt 1 fs",
                    synth.make_protein(2500, 1, radius=synth.compact_radius(2500)).system_dict(), 10000, 20)]
    systems = [p.system_dict() for p in prots]
    sizes = np.array([p.n_atoms for p in prots], np.int64)
    pairs_per_step = float(np.sum(sizes * (sizes - 1) / 2))
    log(f"generated {len(systems)} systems, {sizes.sum()} atoms, {pairs_per_step:.3e} pairs/step "
        f"in {time.time() - t0:.1f}s")

    import torch

    from prolix_amd.engine import GBEngine, Langevin
    from prolix_amd.minimize import lbfgs, minimize

    # 2. engine, relaxation, thermostat start
    eng = GBEngine(systems, cutoff=0.0)
    t0 = time.time()
    x = minimize(eng, [s["positions"] for s in systems], steps=a.min_steps)
    if a.lbfgs_steps > 0:
        x = lbfgs(eng, x, steps=a.lbfgs_steps)
    torch.cuda.synchronize()
    log(f"minimised {a.min_steps} steepest-descent + {a.lbfgs_steps} L-BFGS steps in {time.time() - t0:.1f}s")
    params = [Langevin(DT_FS, float(temps[k]), 1.0, seed=1000 + k, safety_caps=True) for k in range(len(systems))]
    eng.init(x, params)
    eng.steps(a.warmup)
    torch.cuda.synchronize()

    # 3. timed region
    t0 = time.perf_counter()
    eng.steps(a.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    sps = a.steps / wall
    value = ns_per_day(sps) * len(systems)
    log(f"timed {a.steps} steps: {wall * 1e3 / a.steps:.3f} ms/step, {value:.1f} aggregate ns/day")
    ke = eng.kinetic().cpu().numpy()
    temp = 2.0 * ke / (3.0 * sizes * 0.0019872041)
    warn = np.sum([eng.warnings(k) for k in range(len(systems))], axis=0)

    # 4. per-kernel HIP-event timings and the roofline of the dominant pair kernel:
    # events inside the step graph over 50 more integrator steps (the timed
    # region's graph with an event pair around every pass)
    prof, timing = pass_timings(eng, 50)
    pair_pass = max(FLOPS_PER_PAIR, key=lambda k: prof[k])
    t_k = prof[pair_pass] * 1e-3
    achieved = FLOPS_PER_PAIR[pair_pass] * pairs_per_step / t_k / 1e12
    pair_total_s = sum(prof[k] for k in FLOPS_PER_PAIR) * 1e-3
    full_name = {"born_pairs": "prx::k_born_pairs(prx::DeviceView)",
                 "force_pairs": "void prx::k_force_pairs<false>(prx::DeviceView)",
                 "chain_pairs": "prx::k_chain_pairs(prx::DeviceView)"}[pair_pass]
    traffic = None
    traffic_file = os.path.join(ROOT, "profiles", "hbm_traffic_batch256.json")
    if os.path.exists(traffic_file) and a.n_systems == 256:
        with open(traffic_file) as fh:
            traffic = json.load(fh)["kernels"].get(full_name, {}).get("bytes")
    roofline = {
        "bound": "valu",
        "kernel": {"born_pairs": "k_born_pairs", "force_pairs": "k_force_pairs<false>",
                   "chain_pairs": "k_chain_pairs"}[pair_pass],
        "achieved": achieved,
        "peak": PEAK_FP32,
        "unit": "TFLOP/s",
        "frac": achieved / PEAK_FP32,
        "peak_unpacked": PEAK_FP32_UNPACKED,
        "frac_unpacked": achieved / PEAK_FP32_UNPACKED,
        "traffic": traffic,
        "traffic_unit": "bytes per launch (HBM, rocprofv3 FETCH_SIZE + WRITE_SIZE scaled by a 256 MiB copy; "
                        "profiles/hbm_traffic_batch256.md)",
        "flops_per_pair": FLOPS_PER_PAIR[pair_pass],
        "pairs_per_launch": pairs_per_step,
        "kernel_ms": prof[pair_pass],
        "all_pair_passes_tflops": FLOPS_TOTAL * pairs_per_step / pair_total_s / 1e12,
        "pass_ms": prof,
        "timing": timing,
    }
    log(f"passes (ms): {json.dumps({k: round(v, 4) for k, v in prof.items()})}")
    eng.close()
    if a.calibration_copy_mb > 0:
        src = torch.ones(int(a.calibration_copy_mb * 2**20) // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.mul(src, 1.0, out=dst)  # reads and writes calibration_copy_mb MiB
        torch.cuda.synchronize()
        del src, dst

    # 5. the plain path: each system alone in its own engine, same parameters
    plain = None
    if not a.no_plain:
        xs = [np.asarray(v) for v in np.split(x.cpu().numpy(), np.cumsum(sizes)[:-1])]
        tot = 0.0
        t_all = time.time()
        for k, s in enumerate(systems):
            e1 = GBEngine(s, cutoff=0.0)
            e1.init(xs[k], params[k])
            e1.steps(5)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            e1.steps(20)
            torch.cuda.synchronize()
            tot += (time.perf_counter() - t1) / 20.0
            e1.close()
        plain = {"aggregate_ns_per_day": ns_per_day(1.0 / tot) * len(systems), "ms_per_mix_step": tot * 1e3,
                 "batched_speedup": value / (ns_per_day(1.0 / tot) * len(systems))}
        log(f"plain path: {tot * 1e3:.2f} ms per mix step ({time.time() - t_all:.0f}s), "
            f"batched speedup {plain['batched_speedup']:.2f}x")

    # 6. one large system with a cutoff (20,000 atoms, 20 A, neighbour-list GB form)
    large_out = None
    if large is not None:
        from scipy.spatial import cKDTree

        ls = large.system_dict()
        rc = 20.0
        e2 = GBEngine(ls, cutoff=rc, gb_cutoff=True)
        xl = minimize(e2, ls["positions"], steps=a.min_steps)
        if a.lbfgs_steps > 0:
            xl = lbfgs(e2, xl, steps=a.lbfgs_steps)
        e2.init(xl, Langevin(DT_FS, 300.0, 1.0, seed=77, safety_caps=True))
        e2.steps(20)
        torch.cuda.synchronize()
        n_large = 200
        t1 = time.perf_counter()
        e2.steps(n_large)
        torch.cuda.synchronize()
        tl = (time.perf_counter() - t1) / n_large
        xe = e2.state()[0].cpu().numpy()
        n_in = len(cKDTree(xe).query_pairs(rc))
        lp, _ = pass_timings(e2, 20)
        e2.close()
        nl = len(ls["charges"])
        large_out = {"atoms": nl, "cutoff_A": rc, "gb_cutoff": True, "ms_per_step": tl * 1e3,
                     "ns_per_day": ns_per_day(1.0 / tl), "pairs_within_cutoff": n_in,
                     "interactions_per_s": n_in / tl, "all_pairs_equivalent_per_s": nl * (nl - 1) / 2 / tl,
                     "pass_ms": lp}
        log(f"large cutoff system: {tl * 1e3:.3f} ms/step, {n_in} pairs within {rc} A")

    # 7. serving replay: Poisson arrivals into the continuous-batching loop
    serving = None
    if serve_prots:
        from prolix_amd.serving import poisson_trace, serve

        reqs = poisson_trace([p.system_dict() for p in serve_prots], rate_per_s=50.0, steps_range=(1000, 50000),
                             seed=6)
        t1 = time.time()
        results, st = serve(reqs, hbm_budget_bytes=200e9, chunk_steps=200)
        # a seeded 5% sample (at least 2) of the served requests: the device's
        # energies and forces at their final positions, checked against the
        # float64 restatement in the CPU leg
        rng_s = np.random.default_rng(17)
        pick = sorted(rng_s.choice(len(results), size=max(2, len(results) // 20), replace=False).tolist())
        served_sample = []
        for k in pick:
            r = results[k]
            sysd = reqs[r.rid].system
            e1 = GBEngine(sysd, cutoff=0.0)
            e_dev, f_dev, _ = e1.energy_forces(r.positions, energies=True)
            served_sample.append((r.rid, sysd, np.asarray(r.positions, np.float32), e_dev.cpu().numpy()[0],
                                  f_dev.cpu().numpy()))
            e1.close()
        offered_rate = 50.0 * st["offered_pair_steps"] / len(reqs)
        serving = {"requests": st["requests"], "rate_per_s": 50.0, "latency_p50_s": st["latency_p50_s"],
                   "latency_p99_s": st["latency_p99_s"], "aggregate_ns_per_day": st["aggregate_ns_per_day"],
                   "pair_interactions_per_s": st["pair_interactions_per_s"], "wall_s": st["wall_s"],
                   "offered_pair_steps_per_s": offered_rate,
                   "trace": f"{len(reqs)} requests, Poisson 50/s, N log-uniform [200, 8000], "
                            "n_steps U[1000, 50000], T U(280, 350) K, HBM budget 200 GB",
                   "failed_requests": st["failed"],
                   "accuracy": None}
        serving["host_time_s"] = st["host_time_s"]
        serving["device_bytes"] = st["device_bytes"]
        log(f"serving: {len(reqs)} requests in {time.time() - t1:.1f}s, p50 {st['latency_p50_s']:.2f}s "
            f"p99 {st['latency_p99_s']:.2f}s, host {json.dumps({k: round(v, 2) for k, v in st['host_time_s'].items()})}")

        # the same size distribution at a load the device sustains: fewer
        # steps per request and a lower arrival rate, chunks of about 50 ms
        reqs = poisson_trace([p.system_dict() for p in sustained_prots], rate_per_s=a.sustained_rate,
                             steps_range=(1000, 5000), seed=9)
        span = reqs[-1].arrival
        results, st = serve(reqs, hbm_budget_bytes=200e9, chunk_steps=1000, chunk_s=0.05)
        offered = st["offered_pair_steps"] / max(span, 1e-9)
        sustained = {"requests": st["requests"], "rate_per_s": a.sustained_rate,
                     "trace": f"{len(reqs)} requests, Poisson {a.sustained_rate:g}/s, N log-uniform [200, 8000], "
                              "n_steps U[1000, 5000], T U(280, 350) K; chunks sized to about 50 ms",
                     "latency_p50_s": st["latency_p50_s"], "latency_p99_s": st["latency_p99_s"],
                     "offered_pair_steps_per_s": offered, "achieved_pair_steps_per_s": st["pair_interactions_per_s"],
                     "aggregate_ns_per_day": st["aggregate_ns_per_day"], "wall_s": st["wall_s"],
                     "chunks": st["chunks"], "failed_requests": st["failed"], "host_time_s": st["host_time_s"],
                     "device_bytes": st["device_bytes"]}
        log(f"sustained serving: offered {offered:.3e} pair-steps/s, achieved {st['pair_interactions_per_s']:.3e}, "
            f"p50 {st['latency_p50_s']:.3f}s p99 {st['latency_p99_s']:.3f}s, "
            f"host {json.dumps({k: round(v, 2) for k, v in st['host_time_s'].items()})}")

        # the batched mix itself served: each system twice (512 requests, all
        # queued at the start, n_steps U[1000, 3000]) with at most 256 live,
        # so finished systems are replaced while the others run
        rng_m = np.random.default_rng(10)
        xs_mix = np.split(x.cpu().numpy(), np.cumsum(sizes)[:-1])
        order = rng_m.permutation(2 * len(systems)) % len(systems)
        from prolix_amd.serving import Request

        mreqs = []
        for q, k in enumerate(order):
            sd = dict(systems[k])
            sd["positions"] = xs_mix[k]
            mreqs.append(Request(sd, int(rng_m.integers(1000, 3001)), float(temps[k]), 5000 + q, 0.0, DT_FS, 1.0, q))
        results, st = serve(mreqs, hbm_budget_bytes=200e9, max_systems=len(systems), chunk_steps=200)
        ul = st["until_last_admission"]
        mix_served = {"requests": st["requests"], "max_live": len(systems),
                      "trace": f"the {len(systems)}-system mix at its relaxed positions, each system twice in a "
                               "seeded order, all queued at t = 0, n_steps U[1000, 3000], chunks of 200 steps",
                      "aggregate_ns_per_day": st["aggregate_ns_per_day"],
                      "aggregate_ns_per_day_until_last_admission": ul["aggregate_ns_per_day"],
                      "batched_ns_per_day": value,
                      "served_over_batched": ul["aggregate_ns_per_day"] / value,
                      "wall_s": st["wall_s"], "until_last_admission_s": ul["wall_s"], "chunks": st["chunks"],
                      "failed_requests": st["failed"], "host_time_s": st["host_time_s"]}
        log(f"served mix: {ul['aggregate_ns_per_day']:.1f} ns/day until the last admission "
            f"({mix_served['served_over_batched']:.3f} of batched), {st['aggregate_ns_per_day']:.1f} over the whole "
            f"run, host {json.dumps({k: round(v, 2) for k, v in st['host_time_s'].items()})}")
        serving["sustained"] = sustained
        serving["mix"] = mix_served

    # 8. one system alone (the 138-atom peptide, the 2,500-atom globule), GPU and CPU
    single_out, single_x = [], []
    for label, sd, gsteps, csteps in singles:
        r, xs = single_system(sd, gsteps, label)
        single_out.append(r)
        single_x.append(xs)
        log(f"single system {r['atoms']} atoms: {r['gpu_ms_per_step']:.4f} ms/step GPU (default path); by mode "
            f"{json.dumps({k: round(v, 4) for k, v in r['ms_per_step_by_step_mode'].items()})}")

    # 9. CPU baseline (bounded sample)
    cpu = None
    if not a.no_cpu:
        t0 = time.time()
        cpu = cpu_baseline(systems, sizes.astype(np.float64))
        log(f"cpu baseline {cpu['value']:.3f} ns/day ({time.time() - t0:.0f}s)")
        # the served sample against the float64 restatement (the checker)
        if serving is not None:
            from oracle import cpu as C

            def f32(d):  # the device's float32 parameters, held in float64 for the checker
                return {k: (np.asarray(v, np.float32).astype(np.float64)
                            if isinstance(v, np.ndarray) and v.dtype.kind == "f" else v) for k, v in d.items()}

            errs = []
            for rid, sysd, xs, e_dev, f_dev in served_sample:
                eo, fo = C.CpuSystem(f32(sysd)).energy_forces(xs.astype(np.float64), np.float64)
                errs.append({"request": rid, "atoms": len(xs), "energy_rel": float(abs(e_dev[-1] - eo[11]) / abs(eo[11])),
                             "force_rms_rel": float(np.sqrt(np.mean((f_dev - fo) ** 2) / np.mean(fo**2)))})
            serving["accuracy"] = {
                "sample": f"{len(errs)} of {serving['requests']} served requests (seeded 5%, at least 2), at their "
                          "final positions, device vs the float64 C++ restatement",
                "energy_rel_max": max(e["energy_rel"] for e in errs),
                "force_rms_rel_max": max(e["force_rms_rel"] for e in errs),
                "per_request": errs}
            log(f"served sample vs float64: energy rel <= {serving['accuracy']['energy_rel_max']:.2e}, "
                f"force rms rel <= {serving['accuracy']['force_rms_rel_max']:.2e}")
        # the same restatement beside each one-system GPU leg, from the same start
        for r, (label, sd, gsteps, csteps), xs in zip(single_out, singles, single_x):
            r.update(cpu_single(sd, xs, csteps, gsteps))
            r["gpu_over_cpu"] = r["cpu_ms_per_step"] / r["gpu_ms_per_step"]
            log(f"single system {r['atoms']} atoms: {r['cpu_ms_per_step']:.2f} ms/step CPU")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "precision": "fp32 pair arithmetic; fixed-order fp64 partial-sum gathers; fp64 Born finish and "
                     "energy / kinetic-energy reductions",
        "data": "synthetic",
        "config": {"workload": "batched mix: 256 synthetic proteins, N log-uniform [300, 5000], no cutoff, "
                               "BAOAB dt 1 fs, gamma 1/ps, T U(280,350) K",
                   "n_systems": len(systems), "atoms": int(sizes.sum()), "pairs_per_step": pairs_per_step,
                   "relaxation": f"{a.min_steps} steepest-descent + {a.lbfgs_steps} L-BFGS steps before dynamics"},
        "pair_interactions_per_s": pairs_per_step * sps,
        "latency_p50_s": serving["latency_p50_s"] if serving else None,
        "latency_p99_s": serving["latency_p99_s"] if serving else None,
        "latency_sustained_p50_s": serving["sustained"]["latency_p50_s"] if serving else None,
        "latency_sustained_p99_s": serving["sustained"]["latency_p99_s"] if serving else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "plain_path": plain,
        "serving": serving,
        "large_cutoff": large_out,
        "single_systems": single_out or None,
        "temperature_K": {"mean": float(np.mean(temp)), "median": float(np.median(temp)), "min": float(np.min(temp)),
                          "max": float(np.max(temp)), "systems_above_1000K": int(np.sum(temp > 1000.0)),
                          "median_over_target": float(np.median(temp / temps))},
        "safety_cap_events": {"vlimit": int(warn[0]), "force_cap": int(warn[1]), "dx_cap": int(warn[3])},
    }
    line = json.dumps(out)
    print(line, flush=True)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(line + "\n")


if __name__ == "__main__":
    main()
