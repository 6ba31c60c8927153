#!/usr/bin/env python3
"""Particle-turns/s of the fused longitudinal turn loop on one MI355X.

Workload (default ``ring3_1e8``): the 3 GeV ring (C = 234 m, h = 372,
V_rf = 3.2 MV, U0 = 1 MeV/turn, alpha_c = 1.1e-3), 1e8 FP64 macroparticles
(z ~ N(0, 4 mm), dE ~ N(0, 3 MeV), seed 11), 16384 bins over +-8 sigma_z,
two broadband resonators (10 kOhm, 5 GHz, Q=1) + (2 kOhm, 1.2 GHz, Q=5),
10 nC, all five physics flags on.  A step is one turn over the whole bunch;
the bunch is generated on the device (synthetic data) and stays in HBM.

Output: one JSON line.  ``roofline`` prices the dominant kernel (the turn
pass) at 32 algorithmic bytes per particle-turn (read and write z and dE)
against the 8 TB/s HBM peak, with its duration from HIP events on the
library's stream; ``cpu_baseline`` is the oracle (the C restatement of the
reference loop, OpenMP rank-equivalents) timed on the host cores with the
reference's method (median of fresh-copy runs after a warm-up turn) on the
same bunch, bins and wake, at P = all available cores (up to 16) and P = 1.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import struct
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_PARTICLE_TURN = 32.0
METRIC = "particle-turns/s at 1 GPU (% HBM roofline); scan p50/p99 latency vs plain"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def tracking_options(wl: dict, seed: int, timing: bool = False):
    import haissinskijulia_amd as H
    return H.TrackingOptions(seed=seed, resonators=wl["resonators"], charge=wl["charge"],
                             bin_half_width_sigmas=wl["half_width"], timing=timing)


def profiled_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (None if there is none)."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1]), os.path.relpath(path, ROOT)
    except OSError:
        pass
    return None, None


def host_cpu() -> tuple[str, int, int]:
    """(CPU model, physical cores of the host, logical CPUs this process may use)."""
    model, phys = "unknown", set()
    try:
        with open("/proc/cpuinfo") as f:
            pid = cid = None
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name":
                    model = v
                elif k == "physical id":
                    pid = v
                elif k == "core id":
                    cid = v
                elif not k and pid is not None:
                    phys.add((pid, cid))
                    pid = cid = None
            if pid is not None:
                phys.add((pid, cid))
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, len(phys) or (os.cpu_count() or 1), avail


def cpu_baseline(wl: dict, z0, e0, samples_all: int, turns_all: int, samples_one: int,
                 turns_one: int) -> dict:
    """The reference's timing method (tests/benchmark_total.jl:347-375): a
    one-turn warm-up, then the median of several runs, each from a fresh copy
    of the initial bunch, on the full workload shape (the bunch, bins and
    wake the GPU runs), at P = the cores this process may use (up to 16,
    the GPU's host share) and at P = 1.  The timed code is the oracle, the C
    restatement of the reference loop (MPI branch, one OpenMP thread per rank
    equivalent), not the Julia reference, which is not installed."""
    import numpy as np
    from haissinskijulia_amd import workloads as W
    from oracle import oracle as O
    model, phys, avail = host_cpu()
    cores = max(1, min(16, avail))
    n = z0.size
    kw = dict(resonators=wl["resonators"], charge=wl["charge"], bin_half_width_sigmas=wl["half_width"])

    def timed(turns, ranks):
        p = W.ring_3gev_lattice(n_turns=turns)
        t0 = time.perf_counter()
        _, _, r = O.evolve(z0, e0, wl["nbins"], p, seed=wl["seed"], nranks=ranks, **kw)
        dt = time.perf_counter() - t0
        assert r.status == 0
        return dt

    timed(1, cores)   # warm-up turn (benchmark_total.jl:265-271)
    t_all = [timed(turns_all, cores) for _ in range(samples_all)]
    log(f"cpu baseline: P={cores}: {[round(t, 2) for t in t_all]} s for {turns_all} turns")
    t_one = [timed(turns_one, 1) for _ in range(samples_one)] if samples_one > 0 else []
    if t_one:
        log(f"cpu baseline: P=1: {[round(t, 2) for t in t_one]} s for {turns_one} turn(s)")
    med_all = float(np.median(t_all))
    out = {"value": n * turns_all / med_all, "unit": "particle-turns/s", "cores": cores,
           "kind": "port",
           "sample": f"the full workload shape ({n:.0e} particles, {wl['nbins']} bins, the same "
                     f"wake and lattice): median of {samples_all} runs of {turns_all} turns at "
                     f"P = {cores}, each from a fresh copy of the initial bunch after a one-turn "
                     f"warm-up (tests/benchmark_total.jl:347-375); oracle C restatement of the "
                     f"reference loop (MPI branch), not the Julia reference; each run's time "
                     f"includes its set-up (initial spreads, tables), as the reference's does",
           "cpu_model": model, "host_physical_cores": phys, "host_cpus_available": avail}
    if t_one:
        med_one = float(np.median(t_one))
        out["p1_value"] = n * turns_one / med_one
        out["p1_sample"] = f"median of {samples_one} runs of {turns_one} turn(s) at P = 1"
    return out


def cpu_beside(H, W, dev, scan_gpu) -> list:
    """The oracle beside the GPU path on the other workloads' shapes, each a
    one-turn warm-up and one timed run at P = the cores this process may use
    (up to 16): ring3_1e5 in full (1000 turns), ring3_1e7 truncated to 2
    turns (per-turn rates), and the 256-bunch scan truncated to 10 turns and
    run as 256 plain runs in sequence.  Each row also gives the GPU's rate on
    the same shape from this run (the two single bunches timed here, the scan
    from its own leg).  The timed code is the oracle, the C restatement of
    the reference loop, not the Julia reference."""
    import numpy as np
    from oracle import oracle as O
    _, _, avail = host_cpu()
    cores = max(1, min(16, avail))
    rows = []
    for name, t_cpu, t_gpu in (("ring3_1e5", 1000, 1000), ("ring3_1e7", 2, 200)):
        wl = W.WORKLOADS[name]
        n, nbins = wl["n"], wl["nbins"]
        z0, e0 = initial_bunch(H, W, dev, wl)
        kw = dict(resonators=wl["resonators"], charge=wl["charge"],
                  bin_half_width_sigmas=wl["half_width"])
        O.evolve(z0, e0, nbins, W.ring_3gev_lattice(n_turns=1), seed=wl["seed"], nranks=cores, **kw)
        t0 = time.perf_counter()
        _, _, r = O.evolve(z0, e0, nbins, W.ring_3gev_lattice(n_turns=t_cpu), seed=wl["seed"],
                           nranks=cores, **kw)
        dt_cpu = time.perf_counter() - t0
        assert r.status == 0
        b = H.DeviceBunch(n, dev)
        b.upload(z0, e0)
        b.evolve(H.SimulationParameters.from_dict(W.ring_3gev_lattice(n_turns=20)), nbins,
                 tracking_options(wl, wl["seed"]))
        b.upload(z0, e0)
        t0 = time.perf_counter()
        b.evolve(H.SimulationParameters.from_dict(W.ring_3gev_lattice(n_turns=t_gpu)), nbins,
                 tracking_options(wl, wl["seed"]))
        dt_gpu = time.perf_counter() - t0
        b.close()
        rows.append({"workload": f"{name}: {n:.0e} particles, {nbins} bins",
                     "cpu_particle_turns_per_s": n * t_cpu / dt_cpu, "cpu_turns": t_cpu,
                     "gpu_particle_turns_per_s": n * t_gpu / dt_gpu, "gpu_turns": t_gpu,
                     "cores": cores})
        log(f"cpu beside: {name}: oracle {n * t_cpu / dt_cpu:.3e}, GPU {n * t_gpu / dt_gpu:.3e} "
            f"particle-turns/s")
    specs = W.scan_bunches()
    t_cpu, total, dt_cpu = 10, 0, 0.0
    for sp in specs:   # a Gaussian bunch of the bunch's size (the timing needs no more)
        rng = np.random.default_rng(sp["seed"])
        z0 = rng.normal(0.0, W.SIGMA_Z, sp["n"])
        e0 = rng.normal(0.0, W.SIGMA_E, sp["n"])
        kw = dict(resonators=sp["resonators"], charge=sp["charge"],
                  bin_half_width_sigmas=sp["half_width"])
        t0 = time.perf_counter()
        _, _, r = O.evolve(z0, e0, sp["nbins"], W.ring_3gev_lattice(t_cpu, voltage=sp["V"]),
                           seed=sp["seed"] + 1, nranks=cores, **kw)
        dt_cpu += time.perf_counter() - t0
        total += sp["n"] * t_cpu
    rows.append({"workload": f"the 256-bunch scan ({sum(sp['n'] for sp in specs):.3e} particles, "
                             f"4096 bins), 256 plain runs in sequence",
                 "cpu_particle_turns_per_s": total / dt_cpu, "cpu_turns": t_cpu,
                 "gpu_particle_turns_per_s": scan_gpu, "gpu_turns": None, "cores": cores})
    log(f"cpu beside: scan: oracle {total / dt_cpu:.3e} particle-turns/s (GPU batched "
        f"{scan_gpu if scan_gpu is None else f'{scan_gpu:.3e}'})")
    return rows


def _pct(x, q):
    import numpy as np
    return float(np.percentile(np.asarray(x, dtype=float), q)) if len(x) else None


def bench_scan(dev, turns: int, with_plain: bool = True) -> dict:
    """256 uneven bunches (a parameter scan) advanced together by
    hm_batch_evolve, against the plain path: the same bunches evolved one
    after another (skipped with ``with_plain`` false, e.g. under a profiler).
    All bunches are submitted at t = 0, so a bunch's latency is its
    completion time."""
    import haissinskijulia_amd as H
    from haissinskijulia_amd import workloads as W
    specs = W.scan_bunches()

    def params(s):
        return H.SimulationParameters.from_dict(W.ring_3gev_lattice(turns, voltage=s["V"]))

    def opts(s):
        return H.TrackingOptions(seed=s["seed"] + 1, resonators=s["resonators"], charge=s["charge"],
                                 bin_half_width_sigmas=s["half_width"])

    plain, done = [], 0.0
    for s in (specs if with_plain else []):
        b = H.DeviceBunch(s["n"], dev)
        b.generate(s["seed"], W.SIGMA_Z, W.SIGMA_E)
        t0 = time.perf_counter()
        r = b.evolve(params(s), s["nbins"], opts(s))
        done += time.perf_counter() - t0
        plain.append((done, r))
        b.close()
    bunches = []
    for s in specs:
        b = H.DeviceBunch(s["n"], dev)
        b.generate(s["seed"], W.SIGMA_Z, W.SIGMA_E)
        bunches.append(b)
    t0 = time.perf_counter()
    rs = H.batch_evolve(bunches, [s["nbins"] for s in specs], [params(s) for s in specs],
                        [opts(s) for s in specs])
    t_batch = time.perf_counter() - t0
    for b in bunches:
        b.close()
    # bitwise: an unstable parameter point ends in NaN on both paths, and NaN
    # never compares equal
    def bits(r):
        return struct.pack("<3d", r.sigma_E, r.sigma_z, r.E0)
    nonfinite = sum(1 for a in rs if not all(math.isfinite(x) for x in (a.sigma_E, a.sigma_z, a.E0)))
    work = sum(s["n"] for s in specs) * turns
    out = {"workload": f"{len(specs)} bunches, N log-uniform [1e4, 4e6] (sum {work // turns:.3g}), "
                       f"V U[2.5, 4.0] MV, q U[0.5, 10] nC, 4096 bins, {turns} turns, "
                       "seeds from SeedSequence(3)",
           "batched_s": t_batch, "batched_particle_turns_per_s": work / t_batch,
           "p50_s": t_batch, "p99_s": t_batch, "bunches_ending_non_finite": nonfinite}
    if plain:
        lat_plain = [t for t, _ in plain]
        out.update({"plain_s": lat_plain[-1], "speedup": lat_plain[-1] / t_batch,
                    "plain_particle_turns_per_s": work / lat_plain[-1],
                    "plain_p50_s": _pct(lat_plain, 50), "plain_p99_s": _pct(lat_plain, 99),
                    "results_identical_to_plain": all(bits(a) == bits(b)
                                                      for a, (_, b) in zip(rs, plain))})
    return out


def bench_serve(dev, count: int, rate: float, turns_per_tick: int, n_max: float = 1e6,
                turn_range=(100, 2000)) -> dict:
    """Continuous batching (hm_server) of a Poisson job stream in real time,
    against the plain path serving the same stream one job at a time (FIFO;
    its latencies follow from the measured service times and the arrivals)."""
    import haissinskijulia_amd as H
    from haissinskijulia_amd import workloads as W
    jobs = W.serve_jobs(count, rate, n_range=(1e4, n_max), turn_range=turn_range)

    def job(j):
        return H.Job(params=H.SimulationParameters.from_dict(
            W.ring_3gev_lattice(j["turns"], voltage=j["V"])), n=j["n"], nbins=j["nbins"],
            particle_seed=j["seed"], sigma_z=W.SIGMA_Z, sigma_E=W.SIGMA_E,
            options=H.TrackingOptions(seed=j["seed"] + 1, resonators=j["resonators"],
                                      charge=j["charge"], bin_half_width_sigmas=j["half_width"]))

    # plain path: service time of each job alone
    service = []
    t_log = time.perf_counter()
    for i, j in enumerate(jobs):
        if time.perf_counter() - t_log > 30:
            log(f"serving: plain {i}/{len(jobs)} jobs done")
            t_log = time.perf_counter()
        b = H.DeviceBunch(j["n"], dev)
        t0 = time.perf_counter()
        b.generate(j["seed"], W.SIGMA_Z, W.SIGMA_E)
        spec = job(j)
        b.evolve(spec.params, spec.nbins, spec.options)
        service.append(time.perf_counter() - t0)
        b.close()
    lat_plain, free = [], 0.0
    for j, s in zip(jobs, service):
        free = max(free, j["arrival"]) + s
        lat_plain.append(free - j["arrival"])
    # served, in real time
    srv = H.Server(budget_bytes=200 << 30, max_slots=64, device=dev)
    t0 = time.perf_counter()
    submitted, lat, k = {}, [], 0
    while len(lat) < len(jobs):
        now = time.perf_counter() - t0
        while k < len(jobs) and jobs[k]["arrival"] <= now:
            submitted[srv.submit(job(jobs[k]))] = jobs[k]["arrival"]
            k += 1
        active = srv.tick(turns_per_tick)
        tnow = time.perf_counter() - t0
        if time.perf_counter() - t_log > 30:
            log(f"serving: served {len(lat)}/{len(jobs)} jobs, {k} submitted, t = {tnow:.0f} s")
            t_log = time.perf_counter()
        for r in srv.poll():
            assert r.status == 0, r.status
            lat.append(tnow - submitted[r.job_id])
        if active == 0 and k < len(jobs):
            time.sleep(max(0.0, jobs[k]["arrival"] - (time.perf_counter() - t0)))
    span = time.perf_counter() - t0
    srv.close()
    work = sum(j["n"] * j["turns"] for j in jobs)
    span_plain = max(free, 0.0)
    return {"workload": f"{count} jobs, Poisson {rate:g}/s, N log-uniform [1e4, {n_max:.0e}], "
                        f"turns log-uniform [{turn_range[0]}, {turn_range[1]}], V U[2.5, 4.0] MV, "
                        f"q U[0.5, 10] nC, 20% FD pairs (V x (1 -+ 0.5%), common random "
                        f"numbers), 1024 bins, {turns_per_tick} turns per tick, 200 GB HBM budget",
            "p50_s": _pct(lat, 50), "p99_s": _pct(lat, 99),
            "particle_turns_per_s": work / span,
            "plain_p50_s": _pct(lat_plain, 50), "plain_p99_s": _pct(lat_plain, 99),
            "plain_particle_turns_per_s": work / span_plain,
            "offered_particle_turns_per_s": work / jobs[-1]["arrival"]}


def bench_main(H, W, dev, args, wl: dict, out: dict):
    """The single-bunch leg: warm-up, the timed evolve call (the JSON line's
    value), a separately timed run for the pass kernel's roofline, and the
    plain-copy ceiling."""
    n, nbins = wl["n"], wl["nbins"]
    bunch = H.DeviceBunch(n, dev)
    bunch.generate(wl["seed"], W.SIGMA_Z, W.SIGMA_E)

    def params(turns):
        return H.SimulationParameters.from_dict(W.ring_3gev_lattice(n_turns=turns))

    if args.warmup > 0:
        bunch.evolve(params(args.warmup), nbins, tracking_options(wl, wl["seed"]))
        log(f"warmup {args.warmup} turns done")

    t0 = time.perf_counter()
    r = bunch.evolve(params(args.steps), nbins, tracking_options(wl, wl["seed"] + 1))
    t1 = time.perf_counter()
    wall = t1 - t0
    value = n * args.steps / wall
    log(f"timed {args.steps} turns: {wall:.3f} s -> {value:.3e} particle-turns/s "
        f"(device {r.ms_total:.1f} ms), sigma_E={r.sigma_E:.4e} sigma_z={r.sigma_z:.4e} E0={r.E0:.6e}")

    rt = bunch.evolve(params(max(1, args.timing_turns)), nbins,
                      tracking_options(wl, wl["seed"] + 2, True))
    pass_ms = rt.ms_pass_avg
    achieved = BYTES_PER_PARTICLE_TURN * n / (pass_ms * 1e-3) / 1e9
    log(f"turn-pass kernel: {pass_ms:.4f} ms avg over {rt.pass_launches} launches -> {achieved:.0f} GB/s; "
        f"convolution {rt.ms_conv_avg:.4f} ms")
    bunch.close()
    copy_ms, copy_gbs = H.api.stream_copy(int(16 * n), 10, dev)
    log(f"plain copy of {16 * n / 1e9:.1f} GB: {copy_ms:.3f} ms -> {copy_gbs:.0f} GB/s")
    traffic, traffic_src = profiled_traffic("k_turn_pass", args.workload)

    out.update({
        "value": value,
        "ms_per_step": wall / args.steps * 1e3,
        "config": {"workload": args.workload, "particles": n, "nbins": nbins, "turns": args.steps,
                   "bin_half_width_sigmas": wl["half_width"], "charge_C": wl["charge"],
                   "wake": "broadband resonators " + ", ".join(f"({a:g} Ohm, {b:g} Hz, Q={c:g})"
                                                           for a, b, c in wl["resonators"]),
                   "flags": "wakefield, update_eta, update_E0, SR_damping, excitation",
                   "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": "k_turn_pass", "kernel_ms": pass_ms,
                     "algorithmic_bytes_per_launch": BYTES_PER_PARTICLE_TURN * n},
        "whole_loop_frac_of_hbm_peak": value * BYTES_PER_PARTICLE_TURN / 1e9 / HBM_PEAK_GBS,
        "stream_copy_GBps": copy_gbs,
        "conv_ms": rt.ms_conv_avg,
    })


def initial_bunch(H, W, dev, wl: dict):
    """The workload's initial bunch on the host (the particles the GPU legs
    start from), for the CPU baseline; fetched after the GPU legs."""
    b = H.DeviceBunch(wl["n"], dev)
    b.generate(wl["seed"], W.SIGMA_Z, W.SIGMA_E)
    z, e = b.download()
    b.close()
    return z, e


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000, help="timed turns")
    ap.add_argument("--warmup", type=int, default=20, help="untimed turns first")
    ap.add_argument("--workload", default="ring3_1e8")
    ap.add_argument("--timing-turns", type=int, default=200,
                    help="turns of the separate run that times each kernel with HIP events")
    ap.add_argument("--cpu-samples", type=int, default=5,
                    help="fresh-copy runs of the CPU baseline at P = all cores (median)")
    ap.add_argument("--cpu-turns", type=int, default=2, help="turns of each such run")
    ap.add_argument("--cpu-samples-p1", type=int, default=3,
                    help="fresh-copy runs at P = 1 (median); 0 skips them")
    ap.add_argument("--cpu-turns-p1", type=int, default=1, help="turns of each P = 1 run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cpu-beside", action="store_true",
                    help="skip the oracle on the other workloads' shapes (about 20 s)")
    ap.add_argument("--scan-turns", type=int, default=5000,
                    help="turns of the 256-bunch scan (batched vs plain, about 80 s at 5000); "
                         "0 skips it")
    ap.add_argument("--scan-batched-only", action="store_true",
                    help="time the scan's batched path only (no plain runs)")
    ap.add_argument("--no-main", action="store_true",
                    help="skip the single-bunch leg (its line then carries no value)")
    ap.add_argument("--serve-jobs", type=int, default=300,
                    help="jobs of the serving run (server vs plain FIFO); 0 skips it")
    ap.add_argument("--serve-rate", type=float, default=150.0, help="job arrivals per second")
    ap.add_argument("--serve-tick", type=int, default=25, help="turns per server tick")
    ap.add_argument("--serve-n-max", type=float, default=1e6,
                    help="largest particle count of a served job (the full stream: 1e7)")
    ap.add_argument("--serve-turn-range", default="100,2000",
                    help="turn counts of served jobs, lo,hi (the full stream: 1000,20000)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU; the turn loop's state of one bunch "
                                   "lives in one GPU's HBM (no multi-GPU path)",
                          "n_gpus": args.gpus}), flush=True)
        return 2

    import haissinskijulia_amd as H
    from haissinskijulia_amd import workloads as W

    wl = W.WORKLOADS[args.workload]
    dev = H.MI355X(0)
    log(f"device {dev.info()}; workload {args.workload}: n={wl['n']} nbins={wl['nbins']}")
    out = {"metric": METRIC, "unit": "particle-turns/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic"}
    if not args.no_main:
        bench_main(H, W, dev, args, wl, out)
    if args.scan_turns > 0:
        log(f"scan: 256 bunches x {args.scan_turns} turns, plain then batched ...")
        out["scan"] = bench_scan(dev, args.scan_turns, not args.scan_batched_only)
        sc = out["scan"]
        if "plain_s" in sc:
            log(f"scan: batched {sc['batched_s']:.2f} s, plain {sc['plain_s']:.2f} s "
                f"-> {sc['speedup']:.1f}x, identical={sc['results_identical_to_plain']}")
        else:
            log(f"scan: batched {sc['batched_s']:.3f} s")
    if args.serve_jobs > 0:
        log(f"serving: {args.serve_jobs} jobs at {args.serve_rate:g}/s, plain then served ...")
        lo, hi = (int(float(x)) for x in args.serve_turn_range.split(","))
        out["serving"] = bench_serve(dev, args.serve_jobs, args.serve_rate, args.serve_tick,
                                     args.serve_n_max, (lo, hi))
        sv = out["serving"]
        log(f"serving: p50 {sv['p50_s']:.3f} s p99 {sv['p99_s']:.3f} s; plain p50 "
            f"{sv['plain_p50_s']:.3f} s p99 {sv['plain_p99_s']:.3f} s")
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(wl, *initial_bunch(H, W, dev, wl), args.cpu_samples,
                                           args.cpu_turns, args.cpu_samples_p1, args.cpu_turns_p1)
        log(f"cpu baseline: {out['cpu_baseline']['value']:.3e} particle-turns/s")
    else:
        out["cpu_baseline"] = None
    if not args.no_cpu_baseline and not args.no_cpu_beside:
        log("cpu beside the GPU on the other workloads (oracle) ...")
        out["cpu_beside"] = cpu_beside(H, W, dev, (out.get("scan") or {}).get(
            "batched_particle_turns_per_s"))
    print(json.dumps(out), flush=True)
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
