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
reference loop, OpenMP rank-equivalents) timed on a bounded sample on the
host cores.
"""
from __future__ import annotations

import argparse
import json
import os
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
                             bin_half_width_sigmas=wl["half_width"], delta_sigma=-1.0,
                             timing=timing)


def cpu_baseline(wl: dict, n: int, turns: int) -> dict:
    from haissinskijulia_amd import workloads as W
    from oracle import oracle as O
    cores = max(1, min(16, os.cpu_count() or 1))
    p = W.ring_3gev_lattice(n_turns=turns)
    z, e = O.generate_gaussian(wl["seed"], 0.0, W.SIGMA_Z, 0.0, W.SIGMA_E, n)
    t0 = time.perf_counter()
    _, _, r = O.evolve(z, e, wl["nbins"], p, seed=wl["seed"], nranks=cores,
                       resonators=wl["resonators"], charge=wl["charge"],
                       bin_half_width_sigmas=wl["half_width"], delta_sigma=-1.0)
    dt = time.perf_counter() - t0
    assert r.status == 0
    return {"value": n * turns / dt, "unit": "particle-turns/s", "cores": cores, "kind": "port",
            "sample": f"{n:.0e} particles x {turns} turns of the same lattice, bins and wake; "
                      f"oracle C restatement of the reference loop (MPI branch), "
                      f"{cores} OpenMP rank-equivalents, {dt:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000, help="timed turns")
    ap.add_argument("--warmup", type=int, default=20, help="untimed turns first")
    ap.add_argument("--workload", default="ring3_1e8")
    ap.add_argument("--timing-turns", type=int, default=200,
                    help="turns of the separate run that times each kernel with HIP events")
    ap.add_argument("--cpu-particles", type=int, default=4_000_000)
    ap.add_argument("--cpu-turns", type=int, default=60)
    ap.add_argument("--no-cpu-baseline", action="store_true")
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
    n, nbins = wl["n"], wl["nbins"]
    dev = H.MI355X(0)
    info = dev.info()
    log(f"device {info}; workload {args.workload}: n={n} nbins={nbins}")
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

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-turns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "particles": n, "nbins": nbins, "turns": args.steps,
                   "bin_half_width_sigmas": wl["half_width"], "charge_C": wl["charge"],
                   "wake": "broadband resonators " + ", ".join(f"({a:g} Ohm, {b:g} Hz, Q={c:g})"
                                                           for a, b, c in wl["resonators"]),
                   "flags": "wakefield, update_eta, update_E0, SR_damping, excitation",
                   "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_turn_pass", "kernel_ms": pass_ms,
                     "algorithmic_bytes_per_launch": BYTES_PER_PARTICLE_TURN * n},
        "whole_loop_frac_of_hbm_peak": value * BYTES_PER_PARTICLE_TURN / 1e9 / HBM_PEAK_GBS,
        "stream_copy_GBps": copy_gbs,
        "conv_ms": rt.ms_conv_avg,
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_particles, args.cpu_turns)
        log(f"cpu baseline: {out['cpu_baseline']['value']:.3e} particle-turns/s")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
