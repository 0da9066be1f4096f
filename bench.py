"""Benchmark: the Stokes-limit (midpoint) step on one MI355X, 128^3 staggered grid, 65,536
particles at seeded uniform-random positions (random sequential addition: no pair starts
closer than sigma, so no pair sits inside the WCA core), 4-point Peskin kernel, WCA (sigma=1,
rc=2^(1/6), epsilon=kT), kT=1e-3, eta=1, dt=0.05, FP32 fields.

Prints one JSON line: whole-job throughput (cell-updates/s; particle-steps/s alongside), the
roofline of the dominant kernel measured with HIP events over the timed region, the whole-step
algorithmic-byte roofline, and the FP64 oracle timed on the host as the CPU baseline.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure


def algorithmic_bytes(N, npart):
    # per Stokes midpoint step: 192 B/cell (two solves, each: spread write, W + xy-fwd, z pass,
    # xy-inverse, interpolation read, FP32, 3 components) + 200 B/particle
    return 192.0 * N + 200.0 * npart


def kernel_bytes(name, mx, my, mz, npart):
    """Algorithmic HBM bytes of one launch of a kernel group: the per-pass figures of the
    whole-step formula (FP32, 3 components, per Stokes solve), split by the kernel that does the
    pass.  Per cell: spread writes 12 B; the W + xy-forward pass reads 12 and writes 12 (its noise
    is generated in registers); the z pass reads and writes the half spectrum, 24; the xy-inverse
    pass 24; the interpolation reads 12.  Per particle: sort 32, force 24, spread 24,
    interpolation and update 24 (the "wca" group is one launch that also holds the gated sort,
    which runs only when a particle left its tile's skin).  Passes the formula does not charge
    (the spreading records;
    the W pass of the other schemes) get 0 B; their cost is time a fused pipeline would not
    spend.  The Stokes step's noise is drawn in registers inside the z pass."""
    N = mx * my * mz
    per = {
        "spread": 12 * N + 24 * npart,
        "fft_xy_fwd": 24 * N,
        "fft_z_solve": 24 * N,
        "fft_xy_inv": 24 * N,
        "interp": 12 * N + 24 * npart,
        "sort": 32 * npart,
        "wca": 24 * npart,
        "fluid_w": 0,
        "src_prep": 0,
    }
    return per.get(name)


def hbm_traffic(name):
    """Measured HBM bytes per launch of a kernel group, from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE, each scaled by its plain-copy calibration), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("groups", {}).get(name)
    return None if k is None else k.get("hbm_bytes_per_launch")


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _time_oracle(O, p, threads, max_steps, budget_s, all_threads):
    """Seconds per step of the FP64 oracle on `threads` OpenMP threads: one untimed step (on all
    threads), then steps until max_steps or budget_s."""
    O.lib().orc_set_threads(all_threads)
    o = O.Oracle(p)
    o.step(1)
    O.lib().orc_set_threads(threads)
    t0 = time.perf_counter()
    n = 0
    while n < max_steps and (n == 0 or time.perf_counter() - t0 < budget_s):
        o.step(1)
        n += 1
    return (time.perf_counter() - t0) / n, n


def cpu_baseline(args):
    """The FP64 oracle (the reference's discretisation restated on the CPU) timed on this host:
    the benchmarked configuration on all host threads (the reported value) and on one thread,
    and the incompressible 32^3 and inertial-coupling 64^3 configurations likewise.  Bounded
    samples of a few steps each (about 25 s in all)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from fluam_amd._abi import make_params, SCHEME_STOKES_LIMIT, SCHEME_INCOMPRESSIBLE, SCHEME_QUASI_NEUTRALLY_BUOYANT

    all_threads = O.lib().orc_set_threads(0)
    wca = dict(cutoff=2 ** (1 / 6), wca_sigma=1.0)
    p3 = make_params(SCHEME_STOKES_LIMIT, cells=(128, 128, 128), np=65536, kernel_points=4, temperature=1e-3,
                     dt=0.05, seed=3, initial_positions=3, **wca)
    p1 = make_params(SCHEME_INCOMPRESSIBLE, cells=(32, 32, 32), temperature=1e-3, dt=0.05, seed=11)
    p2 = make_params(SCHEME_QUASI_NEUTRALLY_BUOYANT, cells=(64, 64, 64), np=4096, kernel_points=3, temperature=1e-3,
                     dt=0.05, seed=5, initial_positions=3, **wca)
    budget = args.cpu_seconds
    dt_all, n_all = _time_oracle(O, p3, all_threads, 4, 0.25 * budget, all_threads)
    dt_one, n_one = _time_oracle(O, p3, 1, 1, 0.0, all_threads)
    others = {}
    for name, p, steps in (("incompressible 32^3", p1, 40), ("inertial coupling (mass 0) 64^3, 4096 particles", p2, 10)):
        N = p.mx * p.my * p.mz
        a, na = _time_oracle(O, p, all_threads, steps, 0.05 * budget, all_threads)
        s1, n1 = _time_oracle(O, p, 1, max(1, steps // 4), 0.05 * budget, all_threads)
        others[name] = {"cell_updates_per_s_all_threads": N / a, "cell_updates_per_s_1_thread": N / s1,
                        "s_per_step_all_threads": a, "s_per_step_1_thread": s1, "timed_steps": [na, n1]}
    O.lib().orc_set_threads(all_threads)
    N = 128 ** 3
    return {"value": N / dt_all, "unit": "cell-updates/s", "cores": int(all_threads), "kind": "port",
            "particle_steps_per_s": 65536 / dt_all,
            "cpu_model": _cpu_model(), "host_cpus": os.cpu_count(),
            "host_cpus_usable": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None,
            "one_thread": {"value": N / dt_one, "s_per_step": dt_one, "timed_steps": n_one},
            "other_configs": others,
            "sample": f"FP64 oracle, Stokes 128^3 + 65536 particles, {n_all} timed steps after 1 warm-up on "
                      f"{all_threads} OpenMP threads ({dt_all:.2f} s/step) and {n_one} on 1 thread "
                      f"({dt_one:.1f} s/step)"}


def plain_back_to_back(F, params_steps):
    """Seconds to run each job alone, one after another, with eager launches (the baseline the
    batched and served ensembles are compared with): per job, a fresh context, two untimed
    steps, then min(steps, 20) timed steps scaled to the job's step count."""
    total = 0.0
    for p, steps in params_steps:
        s = F.Simulation(p)
        s.step(2)
        s.sync()
        n = min(steps, 20)
        total += s.time_steps(n) / n * steps * 1e-3
        s.close()
    return total


def bench_batched(args):
    """256 independent Stokes-limit replicas, grid edges drawn per axis from {16,24,32,48,64},
    seeded particle counts in [0, 2048] on fluam's default simple-cubic start, 3-point kernel,
    own seeds, graph-replayed steps.  (Uniform-random starts at up to half a particle per cell
    put pairs deep inside the WCA core, whose force then asks for displacements of ~1e16 cells;
    the lattice is the physically sensible start at these densities.)"""
    import numpy as np
    import fluam_amd as F
    from fluam_amd._abi import make_params, SCHEME_STOKES_LIMIT

    rng = np.random.default_rng(4)
    edges = (16, 24, 32, 48, 64)
    params = []
    for i in range(args.replicas):
        shape = tuple(int(rng.choice(edges)) for _ in range(3))
        params.append(make_params(SCHEME_STOKES_LIMIT, cells=shape, np=int(rng.integers(0, 2049)), kernel_points=3,
                                  temperature=1e-3, dt=0.05, seed=1000 + i, initial_positions=1,
                                  cutoff=2 ** (1 / 6), wca_sigma=1.0))
    cells = sum(p.mx * p.my * p.mz for p in params)
    parts = sum(p.np for p in params)
    b = F.Batch(params)
    b.set_mode(args.batch_mode)
    b.step(args.warmup)
    b.sync()
    ms = b.time_steps(args.steps)
    b.sync()
    b.close()
    value = cells * args.steps / (ms * 1e-3)
    plain_s = plain_back_to_back(F, [(p, args.steps) for p in params])
    plain = cells * args.steps / plain_s
    return {
        "metric": "batched-ensemble aggregate cell-updates/s", "value": value, "unit": "cell-updates/s",
        "particle_steps_per_s": parts * args.steps / (ms * 1e-3), "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: seeded grid shapes and particle counts, simple-cubic particle start; zero initial velocity",
        "config": {"workload": f"batched ensemble: {args.replicas} Stokes-limit replicas, edges per axis from "
                               "{16,24,32,48,64}, 0-2048 particles, 3-point kernel, own seeds, graph-replayed steps",
                   "replicas": args.replicas, "total_cells": cells, "total_particles": parts,
                   "batch_mode": ("per-replica graphs", "one graph per batch step, a branch per replica",
                                  "merged stages: one batched dispatch per stage and kernel instance")[args.batch_mode]},
        "back_to_back_plain_cell_updates_per_s": plain,
        "speedup_vs_back_to_back": value / plain,
    }


def bench_config(args):
    """Config 1 (incompressible, 32^3, no particles) or config 2 (inertial coupling, mass 0,
    64^3, 4,096 particles at seeded uniform-random positions, 3-point kernel, WCA), eager
    launches, inputs resident; the FP64 oracle timed on a bounded sample beside it."""
    import fluam_amd as F
    from fluam_amd._abi import make_params, SCHEME_INCOMPRESSIBLE, SCHEME_QUASI_NEUTRALLY_BUOYANT

    if args.workload == "incomp32":
        p = make_params(SCHEME_INCOMPRESSIBLE, cells=(32, 32, 32), temperature=1e-3, dt=0.05, seed=11)
        desc = "schemeIncompressible, 32^3 periodic staggered grid, no particles, kT=1e-3, rho=eta=1, dt=0.05"
    else:
        p = make_params(SCHEME_QUASI_NEUTRALLY_BUOYANT, cells=(64, 64, 64), np=4096, kernel_points=3,
                        temperature=1e-3, dt=0.05, seed=5, initial_positions=3, cutoff=2 ** (1 / 6), wca_sigma=1.0)
        desc = ("schemeQuasiNeutrallyBuoyant (mass 0), 64^3 periodic staggered grid, 4096 particles at seeded "
                "uniform-random positions without pairs closer than sigma, 3-point kernel, WCA sigma=1 "
                "rc=2^(1/6) eps=kT, kT=1e-3, dt=0.05")
    N = p.mx * p.my * p.mz
    sim = F.Simulation(p)
    sim.step(max(args.warmup, 2))
    sim.sync()
    ms = sim.time_steps(args.steps)
    sim.sync()
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        threads = O.lib().orc_set_threads(0)
        o = O.Oracle(p)
        o.step(1)
        t0 = time.perf_counter()
        n = 0
        while time.perf_counter() - t0 < min(args.cpu_seconds, 10.0) and n < 200:
            o.step(1)
            n += 1
        dt = (time.perf_counter() - t0) / n
        cpu = {"value": N / dt, "unit": "cell-updates/s", "cores": int(threads), "kind": "port",
               "sample": f"FP64 oracle, same workload, {n} timed steps after 1 warm-up, OpenMP {threads} threads"}
    return {
        "metric": "fluid cell-updates/s + IB particle-steps/s", "value": N * args.steps / (ms * 1e-3),
        "unit": "cell-updates/s", "particle_steps_per_s": p.np * args.steps / (ms * 1e-3), "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: zero initial velocity, Philox noise" + (", seeded uniform-random positions" if p.np else ""),
        "config": {"workload": desc, "cells": [p.mx, p.my, p.mz], "particles": p.np},
        "cpu_baseline": cpu,
    }


def bench_served(args):
    """Poisson arrivals of Stokes-limit and inertial-coupling (QNB) jobs on cubic grids with edge
    from {32,64,96,128}, seeded particle counts and 50-500 steps, served by the continuous-
    batching loop; p50/p99 latency beside aggregate cell-updates/s."""
    import numpy as np
    import fluam_amd as F
    from fluam_amd import serve
    from fluam_amd._abi import make_params, SCHEME_STOKES_LIMIT, SCHEME_QUASI_NEUTRALLY_BUOYANT

    def make(i, rng):
        edge = int(rng.choice((32, 64, 96, 128)))
        N = edge ** 3
        qnb = bool(rng.integers(0, 2))
        # at most one particle per cell (uniform draw up to 65,536), on fluam's simple-cubic start:
        # uniform-random starts this dense put pairs deep inside the WCA core
        npart = int(rng.integers(0, min(65536, N) + 1))
        p = make_params(SCHEME_QUASI_NEUTRALLY_BUOYANT if qnb else SCHEME_STOKES_LIMIT, cells=(edge, edge, edge),
                        np=npart, kernel_points=3 if qnb else 4, temperature=1e-3, dt=0.05, seed=2000 + i,
                        initial_positions=1, cutoff=2 ** (1 / 6), wca_sigma=1.0)
        return p, int(rng.integers(50, 501))

    jobs = serve.poisson_jobs(args.rate, args.jobs, 5, make)
    res = serve.Server(merged=args.serve_mode == "merged", max_active=args.serve_max_active or None).run(jobs)
    plain_s = plain_back_to_back(F, [(j.params, j.steps) for j in jobs])
    total = sum(j.cell_updates for j in jobs)
    plain = total / plain_s
    return {
        "metric": "served trajectory jobs: aggregate cell-updates/s, p50/p99 job latency",
        "value": res["cell_updates_per_s"], "unit": "cell-updates/s", "n_gpus": 1, "steps": None,
        "warmup": 0, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: seeded job mix and Poisson arrivals",
        "config": {"workload": f"{args.jobs} jobs, Poisson arrivals at {args.rate}/s, cubic edges from "
                               "{32,64,96,128}, Stokes (4-point) or QNB (3-point), 50-500 steps",
                   "jobs": args.jobs, "rate_per_s": args.rate,
                   "serve_mode": ("running jobs regrouped into one merged-stage batch at each join and leave"
                                  if args.serve_mode == "merged" else "a stream and graph replay per job"),
                   "max_active": args.serve_max_active or None},
        "p50_latency_s": res["p50_latency_s"], "p99_latency_s": res["p99_latency_s"],
        "mean_latency_s": res["mean_latency_s"], "wall_s": res["wall_s"],
        "back_to_back_plain_cell_updates_per_s": plain,
        "speedup_vs_back_to_back": res["cell_updates_per_s"] / plain,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--workload", choices=("stokes128", "batched", "served", "incomp32", "qnb64"), default="stokes128",
                    help="stokes128 (default): the headline step; batched / served: the ensemble extensions; "
                         "incomp32 / qnb64: the incompressible and inertial-coupling configurations")
    ap.add_argument("--graph", type=int, default=0, choices=(0, 1),
                    help="stokes128: 1 = replay the steady-state step as a captured HIP graph, 0 = eager launches "
                         "(measured 0.501 vs 0.507 ms per step: the host keeps the queue full either way)")
    ap.add_argument("--replicas", type=int, default=256)
    ap.add_argument("--jobs", type=int, default=24)
    ap.add_argument("--rate", type=float, default=8.0, help="served: job arrivals per second")
    ap.add_argument("--serve-max-active", type=int, default=4,
                    help="served: most jobs running at once (0: as many as the memory budget holds; "
                         "measured on the saturated mix: 1.04x back-to-back at 4, 1.03x at 6, 0.82x unlimited)")
    ap.add_argument("--serve-mode", choices=("streams", "merged"), default="streams",
                    help="served: a stream per job, or the running jobs as one merged-stage batch")
    ap.add_argument("--batch-mode", type=int, default=2, choices=(0, 1, 2),
                    help="batched: 0 = per-replica graphs, 1 = one captured graph per batch step, "
                         "2 = merged stages (one batched dispatch per stage and kernel instance)")
    ap.add_argument("--hw-queues", type=int, default=None,
                    help="hardware queues for the ensemble workloads (HIP reads GPU_MAX_HW_QUEUES at start-up)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the step is a single-GPU kernel pipeline; "
              "how it would shard over several GPUs is a design note, not a run", file=sys.stderr)
        sys.exit(2)
    if args.workload in ("incomp32", "qnb64"):
        if args.steps is None:
            args.steps = 1000 if args.workload == "qnb64" else 100
        print(json.dumps(bench_config(args)))
        return
    if args.workload != "stokes128":
        if args.hw_queues is not None:
            os.environ["GPU_MAX_HW_QUEUES"] = str(max(1, min(args.hw_queues, 16)))
        if args.steps is None:
            args.steps = 200
        line = bench_batched(args) if args.workload == "batched" else bench_served(args)
        print(json.dumps(line))
        return
    if args.steps is None:
        args.steps = 50

    import fluam_amd as F
    from fluam_amd._abi import make_params, SCHEME_STOKES_LIMIT

    mx = my = mz = 128
    npart = 65536
    p = make_params(SCHEME_STOKES_LIMIT, cells=(mx, my, mz), np=npart, kernel_points=4, temperature=1e-3,
                    dt=0.05, seed=3, initial_positions=3, cutoff=2 ** (1 / 6), wca_sigma=1.0)
    sim = F.Simulation(p)
    sim.set_graph(bool(args.graph))  # steady-state steps replay one captured HIP graph
    sim.step(max(args.warmup, 1))
    sim.sync()
    t0 = time.perf_counter()
    ev_ms = sim.time_steps(args.steps)  # HIP events on the context's stream
    sim.sync()
    wall = time.perf_counter() - t0
    ms_per_step = ev_ms / args.steps
    N = mx * my * mz
    cups = N * args.steps / (ev_ms * 1e-3)
    psps = npart * args.steps / (ev_ms * 1e-3)
    # per-kernel times (HIP events around each launch group, same stream) over a short window
    prof = sim.profile(min(args.steps, 10))
    # the dominant kernel: the largest share of the step among the kernels the formula charges
    dom = max((k for k in prof if kernel_bytes(k, mx, my, mz, npart)), key=lambda k: prof[k][0] * prof[k][1])
    kb = kernel_bytes(dom, mx, my, mz, npart)
    kms = prof[dom][0]
    achieved = kb / (kms * 1e-3) / 1e9
    per_kernel = {}
    for k, (k_ms, k_calls) in prof.items():
        b = kernel_bytes(k, mx, my, mz, npart)
        t = hbm_traffic(k)
        per_kernel[k] = {"avg_launch_ms": round(k_ms, 5), "launches_per_step": k_calls / min(args.steps, 10),
                         "algorithmic_bytes": b,
                         "frac": None if not b else round(b / (k_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                         "traffic_bytes": t,
                         "traffic_over_algorithmic": None if (t is None or not b) else round(t / b, 2)}
    step_gbs = algorithmic_bytes(N, npart) / (ms_per_step * 1e-3) / 1e9
    line = {
        "metric": "fluid cell-updates/s + IB particle-steps/s; batched-ensemble p50/p99 latency",
        "value": cups,
        "unit": "cell-updates/s",
        "particle_steps_per_s": psps,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "wall_s": wall,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded uniform-random particle positions without pairs closer than sigma (random "
                "sequential addition), zero initial velocity, Philox noise",
        "config": {"workload": "stokesLimit midpoint step, 128^3 periodic staggered grid, 65536 particles, "
                               "4-point Peskin kernel, WCA sigma=1 rc=2^(1/6) eps=kT, kT=1e-3, eta=1, dt=0.05",
                   "cells": [mx, my, mz], "particles": npart,
                   "initial_positions": "seeded uniform-random by random sequential addition (no pair closer "
                                        "than sigma; data.main initialPositions 3)",
                   "launch": "HIP graph replay of the step" if args.graph else "eager launches",
                   "schedule": sim.schedule()},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": hbm_traffic(dom),
                     "algorithmic_bytes_per_launch": kb, "avg_launch_ms": kms},
        "step_roofline": {"algorithmic_bytes": algorithmic_bytes(N, npart), "achieved_GBs": step_gbs,
                          "frac": step_gbs / HBM_PEAK_GBS},
        "kernels": per_kernel,
        "cpu_baseline": None if args.no_cpu else cpu_baseline(args),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
