"""The headline workload (ring3_1e8) checked against the oracle.

1e8 FP64 particles, z ~ N(0, 4 mm), dE ~ N(0, 3 MeV) from seed 11, 16384
bins over +-8 sigma_z, two broadband resonators (10 kOhm, 5 GHz, Q=1) +
(2 kOhm, 1.2 GHz, Q=5), 10 nC, all five physics flags, the device Philox
stream for the excitation: exactly what the benchmark times, for two turns
(the oracle runs them on the host's cores).  Level 1 of the parity plan:
coordinates to 1e-12 normwise, E0 to 1e-14.  The loop of
src/evolution.jl:295-456 at the benchmarked size, with the partial LDS
window of the pass, the eight-deep interpolation rings and 16 reduction
slices.
"""
import os
import time

import numpy as np
import pytest

from helpers import COORD_RTOL, W, normwise, to_sim_params
from oracle import oracle as O

import haissinskijulia_amd as H

pytestmark = pytest.mark.gpu


def test_ring3_1e8_two_turns_match_oracle():
    wl = W.WORKLOADS["ring3_1e8"]
    n, nbins, turns = wl["n"], wl["nbins"], 2
    p = W.ring_3gev_lattice(turns)
    kw = dict(resonators=wl["resonators"], charge=wl["charge"],
              bin_half_width_sigmas=wl["half_width"])
    t0 = time.time()
    b = H.DeviceBunch(n)
    b.generate(wl["seed"], W.SIGMA_Z, W.SIGMA_E)
    z0, e0 = b.download()
    # the device's initial bunch is the oracle's stream-1 Gaussians of seed 11
    # (the device's Box-Muller rounds its log and sincos differently: an ulp or
    # two); both paths then start from the device's bunch
    zh, eh = O.generate_gaussian(wl["seed"], 0.0, W.SIGMA_Z, 0.0, W.SIGMA_E, 1_000_000)
    np.testing.assert_allclose(z0[:1_000_000], zh, rtol=4e-15, atol=1e-20)
    np.testing.assert_allclose(e0[:1_000_000], eh, rtol=4e-15, atol=1e-5)
    del zh, eh
    print(f"ring3_1e8: bunch generated ({time.time() - t0:.1f} s)", flush=True)
    r = b.evolve(to_sim_params(p), nbins, H.TrackingOptions(seed=wl["seed"] + 1, **kw))
    z, e = b.download()
    b.close()
    print(f"ring3_1e8: GPU {turns} turns done, E0 {r.E0:.15e} ({time.time() - t0:.1f} s)", flush=True)
    ranks = max(1, min(16, os.cpu_count() or 1))
    zr, er, rr = O.evolve(z0, e0, nbins, p, seed=wl["seed"] + 1, nranks=ranks, **kw)
    del z0, e0
    assert rr.status == 0
    ez, ee = normwise(z, zr), normwise(e, er)
    print(f"ring3_1e8: oracle on {ranks} threads done ({time.time() - t0:.1f} s); "
          f"normwise z={ez:.2e} dE={ee:.2e} E0 rel={abs(r.E0 - rr.E0) / rr.E0:.2e}", flush=True)
    assert ez <= COORD_RTOL and ee <= COORD_RTOL
    assert r.E0 == pytest.approx(rr.E0, rel=1e-14)
    assert r.sigma_E == pytest.approx(rr.sigma_E, rel=1e-10)
    assert r.sigma_z == pytest.approx(rr.sigma_z, rel=1e-10)
