"""Numerical Haissinski solution of the reference's turn map (test helper).

With the wake on, update_E0 and update_eta off, SR damping and quantum
excitation on, the map of src/evolution.jl drives a bunch to a stationary
state.  In the smooth-motion limit that state is

    lambda(z) ∝ exp(Phi(z) / (K sigma_E^2)),   Phi(z) = -∫_0^z F(s) ds,
    F(z) = V (sin(phi_s - z rf) - sin phi_s) - P[lambda](z),

with K = coeff / rf the drift per unit energy (evolution.jl:157-164) and
sigma_E the excitation level sigma_E0 (quantum_excitation.jl:58-59 balances
synchrotron_radiation.jl:44-55 at sigma_E = sigma_E0).  The potential P is
the tracker's for a lag-indexed wake with a fixed current (the charge
option, so the fixed point depends on lambda only): counts / n_h on the bin
grid of wakefield.jl:186-319 (the delta kernel belongs to the reference's
dimensionless current and is left out with a charge), each bin feeling the
bins ahead of it (larger z is the head, wakefield.jl:111-113) through the
wake at their distance.  The fixed point is found by damped iteration on the
bin centres.
"""
from __future__ import annotations

import math

import numpy as np

from oracle import oracle as O

C_LIGHT = 299792458.0


def solve(p: dict, nbins: int, z_width: float, resonators, current: float, sigma_E: float,
          iters: int = 400, delta_sigma: float | None = None):
    """Returns (centres, probability per bin, P at the centres)."""
    edges, c, step = O.bin_grid(z_width, nbins)
    wtab = np.array([sum(O.resonator_wake(j * step, *r) for r in resonators)
                     for j in range(nbins)])
    ltab = np.ones(nbins) if delta_sigma is None else np.array([O.delta(x, delta_sigma) for x in c])
    g0 = p["E0"] / p["mass"]
    b0 = math.sqrt(1.0 - 1.0 / (g0 * g0))
    eta0 = p["alpha_c"] - 1.0 / (g0 * g0)
    rf = (p["freq_rf"] * 2.0 * math.pi) / (b0 * C_LIGHT)
    coeff = (2.0 * math.pi * p["harmonic"] * eta0) / (b0 * b0 * p["E0"])
    K = coeff / rf
    f_rf = p["voltage"] * (np.sin(p["phi_s"] - c * rf) - math.sin(p["phi_s"]))
    L = 1 << int(math.ceil(math.log2(2 * nbins)))
    w_hat = np.fft.rfft(np.concatenate([wtab, np.zeros(L - nbins)]))
    i0 = int(np.argmin(np.abs(c)))
    lam = np.exp(-c * c / (2.0 * (0.2 * z_width) ** 2))
    lam /= lam.sum()
    P = np.zeros(nbins)
    for _ in range(iters):
        g = (ltab * lam)[::-1]   # P_k = I sum_j W_j g_{k+j}: the index convolution reversed
        P = current * np.fft.irfft(w_hat * np.fft.rfft(np.concatenate([g, np.zeros(L - nbins)])),
                                   L)[:nbins][::-1]
        F = f_rf - P
        phi = -np.concatenate([[0.0], np.cumsum(0.5 * (F[1:] + F[:-1]) * step)])
        phi -= phi[i0]
        new = np.exp((phi - phi.max()) / (K * sigma_E * sigma_E))
        lam = 0.7 * lam + 0.3 * new / new.sum()
    return c, lam / lam.sum(), P


def moments(c: np.ndarray, prob: np.ndarray):
    """(mean, std, skewness) of a distribution on the bin centres."""
    mu = float((c * prob).sum())
    sd = math.sqrt(float(((c - mu) ** 2 * prob).sum()))
    sk = float(((c - mu) ** 3 * prob).sum()) / sd ** 3
    return mu, sd, sk


def sample_moments(z: np.ndarray):
    mu = float(z.mean())
    sd = float(z.std())
    sk = float(((z - mu) ** 3).mean()) / sd ** 3
    return mu, sd, sk
