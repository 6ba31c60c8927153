"""Lattices and bunch shapes used by the tests, the benchmark and the examples.

Two lattices:

* ``reference_benchmark_lattice`` -- the physics block of HaissinskiJulia's
  integration benchmark (tests/benchmark_total.jl:139-166): 4 GeV, R = 250 m,
  h = 360, V = 5 MV, pipe radius 0.25 mm.
* ``ring_3gev_lattice`` -- the 3 GeV electron ring of the tracking workloads:
  C = 234 m, h = 372, V_rf = 3.2 MV, U0 = 1 MeV/turn (sin phi_s = U0/V),
  alpha_c = 1.1e-3, R = C / 2 pi (the SR damping decrement of
  synchrotron_radiation.jl:44 then gives ~3900 turns), f_rf = h beta c / C.

Parameter dictionaries use the field names and order of
``SimulationParameters`` (src/data_structures.jl:81-97).
"""
from __future__ import annotations

import math

SPEED_LIGHT = 299792458.0          # src/constants.jl:5
ELECTRON_CHARGE = 1.602176634e-19  # src/constants.jl:6
MASS_ELECTRON = 0.51099895069e6    # src/constants.jl:7

PARAM_FIELDS = ("E0", "mass", "voltage", "harmonic", "radius", "pipe_radius", "alpha_c",
                "phi_s", "freq_rf", "n_turns", "use_wakefield", "update_eta", "update_E0",
                "SR_damping", "use_excitation")


def reference_benchmark_lattice(n_turns: int = 100, flags=(True, True, True, True, True)) -> dict:
    """tests/benchmark_total.jl:139-152 (all five physics flags on as at :259)."""
    E0 = 4e9
    mass = MASS_ELECTRON
    gamma = E0 / mass
    beta = math.sqrt(1 - 1 / gamma ** 2)
    phis = 5 * math.pi / 6
    freq_rf = (phis + 10 * math.pi / 180) * beta * SPEED_LIGHT / (2 * math.pi)
    p = dict(E0=E0, mass=mass, voltage=5e6, harmonic=360, radius=250.0, pipe_radius=0.00025,
             alpha_c=3.68e-4, phi_s=phis, freq_rf=freq_rf, n_turns=n_turns)
    for k, v in zip(PARAM_FIELDS[10:], flags):
        p[k] = bool(v)
    return p


def reference_benchmark_bunch_sigmas() -> tuple[float, float]:
    """(sigma_z0, sigma_E0) of tests/benchmark_total.jl:154-166."""
    p = reference_benchmark_lattice()
    E0, alpha_c, h, V = p["E0"], p["alpha_c"], p["harmonic"], p["voltage"]
    gamma = E0 / p["mass"]
    beta = math.sqrt(1 - 1 / gamma ** 2)
    T_rev = (2 * math.pi * p["radius"]) / (beta * SPEED_LIGHT)
    w_rev = 2 * math.pi / T_rev
    sigma_E0 = 1e6
    fac = alpha_c * E0 / (h * V * abs(math.cos(p["phi_s"])))
    sigma_z0 = math.sqrt(2 * math.pi) * SPEED_LIGHT / w_rev * math.sqrt(fac) * sigma_E0 / E0
    return sigma_z0, sigma_E0


def ring_3gev_lattice(n_turns: int = 1000, voltage: float = 3.2e6,
                      flags=(True, True, True, True, True), pipe_radius: float = 0.25e-3) -> dict:
    E0 = 3e9
    mass = MASS_ELECTRON
    circumference = 234.0
    radius = circumference / (2 * math.pi)
    h = 372
    U0 = 1.0e6
    gamma = E0 / mass
    beta = math.sqrt(1 - 1 / gamma ** 2)
    phis = math.pi - math.asin(U0 / voltage)  # above transition: stable for eta > 0
    freq_rf = h * beta * SPEED_LIGHT / circumference
    p = dict(E0=E0, mass=mass, voltage=voltage, harmonic=h, radius=radius,
             pipe_radius=pipe_radius, alpha_c=1.1e-3, phi_s=phis, freq_rf=freq_rf,
             n_turns=n_turns)
    for k, v in zip(PARAM_FIELDS[10:], flags):
        p[k] = bool(v)
    return p


# Broadband resonator of the 3 GeV workloads: (R_s [Ohm], f_r [Hz], Q).
RESONATOR_BB = (10e3, 5e9, 1.0)
RESONATOR_LOWQ = (2e3, 1.2e9, 5.0)

# Bunch shapes: sigma_z = 4 mm, sigma_delta = 1e-3 (sigma_E = 3 MeV at 3 GeV).
SIGMA_Z = 4e-3
SIGMA_E = 3e6

WORKLOADS = {
    # 1e5 particles, 1024 bins over +-6 sigma, 1 nC, 1000 turns, seed 42
    "ring3_1e5": dict(n=100_000, nbins=1024, half_width=6.0, charge=1e-9, turns=1000, seed=42,
                      resonators=[RESONATOR_BB]),
    # 1e7 particles, 8192 bins over +-8 sigma, 5 nC, 10000 turns, seed 7
    "ring3_1e7": dict(n=10_000_000, nbins=8192, half_width=8.0, charge=5e-9, turns=10_000, seed=7,
                      resonators=[RESONATOR_BB]),
    # 1e8 particles (1.6 GB of FP64 state), 16384 bins, 10 nC, two resonators, 5000 turns, seed 11
    "ring3_1e8": dict(n=100_000_000, nbins=16384, half_width=8.0, charge=10e-9, turns=5000,
                      seed=11, resonators=[RESONATOR_BB, RESONATOR_LOWQ]),
}


def scan_bunches(count: int = 256, seed: int = 3) -> list[dict]:
    """The bunches of a batched parameter scan on the 3 GeV ring: particle
    counts log-uniform in [1e4, 4e6], V_rf uniform in [2.5, 4.0] MV, charge
    uniform in [0.5, 10] nC, 4096 bins over +-8 sigma_z, the broadband
    resonator; per-bunch seeds spawned from numpy.random.SeedSequence(seed)."""
    import numpy as np
    ss = np.random.SeedSequence(seed)
    rng = np.random.default_rng(ss)
    n = np.exp(rng.uniform(math.log(1e4), math.log(4e6), count)).astype(np.int64)
    V = rng.uniform(2.5e6, 4.0e6, count)
    q = rng.uniform(0.5e-9, 10e-9, count)
    seeds = [int(c.generate_state(1, np.uint64)[0]) for c in ss.spawn(count)]
    return [dict(n=int(n[i]), V=float(V[i]), charge=float(q[i]), seed=seeds[i], nbins=4096,
                 half_width=8.0, resonators=[RESONATOR_BB]) for i in range(count)]


def serve_jobs(count: int, rate: float, n_range=(1e4, 1e7), turn_range=(1e3, 2e4),
               fd_fraction: float = 0.2, seed: int = 5) -> list[dict]:
    """A stream of tracking jobs on the 3 GeV ring: Poisson arrivals at `rate`
    jobs/s; particle counts and turn counts log-uniform in the given ranges;
    V_rf uniform in [2.5, 4.0] MV and charge uniform in [0.5, 10] nC.  About
    `fd_fraction` of the jobs come as finite-difference pairs (V_rf x (1 -+
    0.5%)) with common random numbers: both members have the same arrival,
    particle seed, QE seed and charge.  1024 bins over +-8 sigma_z, the
    broadband resonator."""
    import numpy as np
    rng = np.random.default_rng(seed)
    jobs: list[dict] = []
    t = 0.0
    while len(jobs) < count:
        t += rng.exponential(1.0 / rate)
        n = int(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1]))))
        turns = int(math.exp(rng.uniform(math.log(turn_range[0]), math.log(turn_range[1]))))
        s = int(rng.integers(1, 2 ** 62))
        V = float(rng.uniform(2.5e6, 4.0e6))
        q = float(rng.uniform(0.5e-9, 10e-9))
        base = dict(arrival=t, n=n, turns=turns, seed=s, V=V, charge=q, nbins=1024,
                    half_width=8.0, resonators=[RESONATOR_BB], pair=-1)
        if len(jobs) + 1 < count and rng.uniform() < fd_fraction / 2:
            k = len(jobs)
            jobs.append(dict(base, V=V * (1 - 0.005), pair=k + 1))
            jobs.append(dict(base, V=V * (1 + 0.005), pair=k))
        else:
            jobs.append(base)
    return jobs
