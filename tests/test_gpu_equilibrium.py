"""Statistical parity: with the device's own random stream, equilibrium
moments and the bunch profile agree with independent oracle ensembles.

Each ensemble member is a 2e4-particle bunch on the 3 GeV ring (all physics
on, broadband resonator, 1 nC) started from z ~ N(0, 4 mm), dE ~ N(0, 3 MeV)
and tracked for 3000 turns (about 0.8 damping times).  GPU members and oracle
members use disjoint seeds, so the two ensembles are independent samples of
the same process; their means must agree within 3 combined standard errors,
and their 32-bin line densities must pass a chi-square test at the 0.1% level.
"""
import numpy as np
import pytest
from scipy import stats

from helpers import W
from oracle import oracle as O

import haissinskijulia_amd as H

pytestmark = pytest.mark.gpu

N = 20_000
TURNS = 3000
NBINS = 512
K_GPU = 24
K_ORACLE = 8
EDGES = np.linspace(-0.024, 0.024, 33)   # +-6 equilibrium sigma_z


def _opts(seed):
    return dict(resonators=[W.RESONATOR_BB], charge=1e-9, bin_half_width_sigmas=8.0, seed=seed)


def _moments(z, e, r):
    prof = np.histogram(z, bins=EDGES)[0] / z.size
    return np.array([r.sigma_E, r.sigma_z, float(np.mean(z)), float(np.mean(e)), r.E0]), prof


def _gpu_member(k):
    p = H.SimulationParameters.from_dict(W.ring_3gev_lattice(TURNS))
    b = H.DeviceBunch(N)
    b.generate(1000 + k, W.SIGMA_Z, W.SIGMA_E)
    r = b.evolve(p, NBINS, H.TrackingOptions(**_opts(5000 + k)))
    z, e = b.download()
    b.close()
    return _moments(z, e, r)


def _oracle_member(k):
    p = W.ring_3gev_lattice(TURNS)
    z0, e0 = O.generate_gaussian(2000 + k, 0.0, W.SIGMA_Z, 0.0, W.SIGMA_E, N)
    z, e, r = O.evolve(z0, e0, NBINS, p, nranks=8, **_opts(6000 + k))
    assert r.status == 0
    return _moments(z, e, r)


def test_equilibrium_moments_and_profile_match_oracle_ensembles():
    g = [_gpu_member(k) for k in range(K_GPU)]
    o = [_oracle_member(k) for k in range(K_ORACLE)]
    gm = np.array([m for m, _ in g])
    om = np.array([m for m, _ in o])
    names = ["sigma_E", "sigma_z", "centroid_z", "centroid_dE", "E0"]
    for i, name in enumerate(names):
        a, b = gm[:, i], om[:, i]
        se = np.sqrt(a.var(ddof=1) / a.size + b.var(ddof=1) / b.size)
        diff = abs(a.mean() - b.mean())
        print(f"{name}: gpu {a.mean():.6e} +- {a.std(ddof=1):.2e}  oracle {b.mean():.6e} +- "
              f"{b.std(ddof=1):.2e}  |diff|/se = {diff / se if se > 0 else 0.0:.2f}", flush=True)
        if se == 0.0:
            assert diff <= 1e-12 * abs(b.mean())
        else:
            assert diff <= 3.0 * se, name
    gp = np.array([p for _, p in g])
    op = np.array([p for _, p in o])
    keep = op.mean(axis=0) > 1e-3          # bins with at least 20 particles on average
    d = gp.mean(axis=0)[keep] - op.mean(axis=0)[keep]
    se2 = gp.var(axis=0, ddof=1)[keep] / K_GPU + op.var(axis=0, ddof=1)[keep] / K_ORACLE
    chi2 = float(np.sum(d * d / se2))
    lim = float(stats.chi2.ppf(0.999, keep.sum()))
    print(f"profile chi2 = {chi2:.1f} over {keep.sum()} bins (0.1% limit {lim:.1f})", flush=True)
    assert chi2 <= lim
