"""The integration benchmark of the turn loop on the GPU, logged in the
reference's format.

``python -m haissinskijulia_amd.benchlog --particles 1e5 --turns 100`` is the
counterpart of ``tests/benchmark_total.jl`` (serial mode): the same lattice
(:139-152), bunch spreads (:154-166), bin count
``next_power_of_two(max(64, round(N/100)))`` (:255), all five physics flags
(:259) and the reference's own wake (pipe radius, no resonators); a one-turn
pre-run (:265-271); then ``samples`` timed evolve calls, each on a fresh copy
of the initial bunch, and their median (:347-360).  The log has the
reference's parameter block (:310-315) and aggregated result lines
(:397-403), so that the reference's log parser (``parse_benchmark_logs.sh``
:62-108) reads processes, threads, particles, turns and time from it.

Differences, by construction: the run mode is ``GPU`` with one process and one
host thread; ``Sum of Median Memory Allocated`` is 0 (the evolve call reuses the
bunch's device workspace after the pre-run); GC time and allocations are 0;
there is no LIKWID section, so the parser's FLOPS columns are NA.
"""
from __future__ import annotations

import argparse
import datetime
import math
import os
import re
import statistics
import sys
import time

from . import workloads as W

# names and order of the parameter block (tests/benchmark_total.jl:310-313)
PARAM_NAMES = ("run_mode", "E0_ini", "E0_start", "mass", "voltage", "harmonic", "radius",
               "pipe_radius", "α_c", "γ", "β", "η", "sin_ϕs", "ϕs", "freq_rf", "μ_z", "μ_E",
               "ω_rev", "σ_E0", "σ_z0", "n_turns", "n_particles_global",
               "num_threads_per_process", "mpi_comm_size", "nbins_calculated")


def format_bytes(nbytes: float) -> str:
    """tests/benchmark_total.jl:58-69: two decimals in the largest binary unit
    below 1024, "0 bytes" for zero."""
    b = int(round(nbytes))
    if b == 0:
        return "0 bytes"
    units = ("bytes", "KiB", "MiB", "GiB", "TiB", "PiB")
    val, idx = float(b), 0
    while val >= 1024 and idx < len(units) - 1:
        val /= 1024
        idx += 1
    return f"{val:.2f} {units[idx]}"


def format_time(t_sec: float) -> str:
    """tests/benchmark_total.jl:71-92, including its unit walk: a time below
    1 s moves to ms, and further to μs and ns only below 1e-6 s and 1e-12 s;
    three decimals."""
    if t_sec == 0:
        return "0.0 s"
    units = ("s", "ms", "μs", "ns")
    scaling = (1.0, 1e3, 1e6, 1e9)
    idx, val = 0, t_sec
    while idx < len(units) - 1 and val * scaling[idx + 1] < 1000:
        idx += 1
        val = t_sec * scaling[idx]
        if idx == len(units) - 1 or val >= 1:
            break
    if idx == 0 and val < 1:
        idx, val = 1, t_sec * scaling[1]
    return f"{val:.3f} {units[idx]}"


def next_power_of_two(n: int) -> int:
    """src/utils.jl:119-122."""
    return 1 << max(0, (int(n) - 1).bit_length())


def log_filename(n_turns: int, n_particles: int, n_gpus: int = 1, threads: int = 1) -> str:
    """turns<T>_particles<N>_gpu<G>_threads<t>.log, with T and N written as
    the reference writes them (tests/benchmark_total.jl:295-302): %.0e, then
    "." and "e+" removed and "e0+" replaced by "e" (1e+05 -> 1e5)."""
    def sci(x):
        return re.sub(r"e0+", "e", f"{float(x):.0e}".replace(".", "").replace("e+", "e"))
    return f"turns{sci(n_turns)}_particles{sci(n_particles)}_gpu{n_gpus}_threads{threads}.log"


def parameters(n_particles: int, n_turns: int) -> dict:
    """The parameter block's values for this run (tests/benchmark_total.jl:139-166,255,286-293)."""
    p = W.reference_benchmark_lattice(n_turns)
    sz, sE = W.reference_benchmark_bunch_sigmas()
    E0, mass = p["E0"], p["mass"]
    gamma = E0 / mass
    beta = math.sqrt(1 - 1 / gamma ** 2)
    T_rev = (2 * math.pi * p["radius"]) / (beta * W.SPEED_LIGHT)
    return {"run_mode": "GPU", "E0_ini": E0, "E0_start": E0, "mass": mass,
            "voltage": p["voltage"], "harmonic": p["harmonic"], "radius": p["radius"],
            "pipe_radius": p["pipe_radius"], "α_c": p["alpha_c"], "γ": gamma, "β": beta,
            "η": p["alpha_c"] - 1 / gamma ** 2, "sin_ϕs": 0.5, "ϕs": p["phi_s"],
            "freq_rf": p["freq_rf"], "μ_z": 0.0, "μ_E": 0.0, "ω_rev": 2 * math.pi / T_rev,
            "σ_E0": sE, "σ_z0": sz, "n_turns": int(n_turns),
            "n_particles_global": int(n_particles), "num_threads_per_process": 1,
            "mpi_comm_size": 1,
            "nbins_calculated": next_power_of_two(max(64, int(round(n_particles / 100))))}


def format_log(params: dict, median_time_s: float, memory_bytes: float = 0.0,
               gc_time_s: float = 0.0, allocations: int = 0, when: datetime.datetime | None = None,
               extra_lines=()) -> str:
    """The log text: header and parameter block (tests/benchmark_total.jl:
    308-315), then the aggregated results (:397-403)."""
    when = when or datetime.datetime.now()
    out = [f"=== Simulation Log ({params['run_mode']} Mode) ===",
           f"Date: {when.strftime('%Y-%m-%d %H:%M:%S')}", "", "=== Parameters ==="]
    out += [f"{k} = {params[k]}" for k in PARAM_NAMES]
    out += ["", "=== Simulation Output ==="]
    out += list(extra_lines)
    gc_pct = (gc_time_s / median_time_s * 100) if median_time_s > 0 else 0.0
    out += ["", "Aggregated Benchmark Results (Based on Local Medians):",
            f"  Max Median Time (across ranks): {format_time(median_time_s)}",
            f"  Sum of Median Memory Allocated: {format_bytes(memory_bytes)}",
            f"  Sum of Median GC Time: {format_time(gc_time_s)} ({gc_pct:.2f}%)",
            f"  Sum of Median Allocations: {allocations:.5g}", "-" * 20]
    return "\n".join(out) + "\n"


def run(n_particles: int, n_turns: int, samples: int = 10, seed: int = 1,
        device=None) -> tuple[dict, list[float], object]:
    """The timed benchmark: returns (parameters, sample times [s], last result)."""
    import numpy as np
    import haissinskijulia_amd as H
    prm = parameters(n_particles, n_turns)
    dev = device or H.MI355X.default()
    nbins = prm["nbins_calculated"]
    lattice = H.SimulationParameters.from_dict(W.reference_benchmark_lattice(n_turns))
    pre = H.SimulationParameters.from_dict(W.reference_benchmark_lattice(1))
    opts = H.TrackingOptions(seed=seed)
    b = H.DeviceBunch(n_particles, dev)
    try:
        # initial bunch: generate_particles (evolution.jl:42-133) with the
        # benchmark's spreads, kept on the host for the fresh copies
        b.generate_particles(seed, prm["μ_z"], prm["μ_E"], prm["σ_z0"], prm["σ_E0"])
        z0, e0 = (np.array(a, copy=True) for a in b.download())
        b.evolve(pre, nbins, opts)                    # pre-run (compilation run)
        times, r = [], None
        for _ in range(samples):
            b.upload(z0, e0)                          # fresh copy: set-up, untimed
            t0 = time.perf_counter()
            r = b.evolve(lattice, nbins, opts)
            times.append(time.perf_counter() - t0)
    finally:
        b.close()
    return prm, times, r


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--particles", type=float, default=1e5)
    ap.add_argument("--turns", type=float, default=1e2)
    ap.add_argument("--samples", type=int, default=10)
    ap.add_argument("--log-dir", default=os.path.join("logs", "gpu_run"))
    args = ap.parse_args(argv)
    n, t = int(args.particles), int(args.turns)
    prm, times, r = run(n, t, args.samples)
    med = statistics.median(times)
    os.makedirs(args.log_dir, exist_ok=True)
    path = os.path.join(args.log_dir, log_filename(t, n))
    extra = [f"Final: sigma_E = {r.sigma_E}, sigma_z = {r.sigma_z}, E0 = {r.E0}",
             f"Samples (s): {', '.join(f'{x:.6f}' for x in times)}"]
    with open(path, "w") as f:
        f.write(format_log(prm, med, extra_lines=extra))
    print(f"{path}: median {format_time(med)} over {len(times)} samples "
          f"({n * t / med:.3e} particle-turns/s)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
