"""Pins of the oracle by what the mathematics fixes, for the four reference
functions whose restatement no vector or brute force covers:

* the energy-dependent slip factor of the drift (src/evolution.jl:330-344),
* the wake function of the pipe-radius model (src/wakefield.jl:115-118),
* the delta line-density kernel (src/utils.jl:58-66),
* the wake current heuristic (src/evolution.jl:373-386).

Each expectation is a closed form, an asymptotic expansion, an integral or a
scaling law, never the oracle's own expression retyped, and each is chosen so
that a plausible misreading fails: gamma0 in place of the particle's gamma,
a missing square, a sign flip of the exponent, a dropped factor.
"""
from __future__ import annotations

import math

import mpmath as mp
import numpy as np
import pytest

from oracle import oracle as O
from haissinskijulia_amd import workloads as W

mp.mp.dps = 40


# --------------------------------------------------------------------------
# drift with update_eta (evolution.jl:330-344)
# --------------------------------------------------------------------------

def _drift_pair(z, dE, mass, gamma0, alpha_c, h, beta0, E0, rf, phis):
    """(update_eta drift, constant-eta drift at eta0 = alpha_c - 1/gamma0^2)."""
    zu = O.drift_update_eta(z, dE, mass, gamma0, alpha_c, h, beta0, E0, rf, phis)
    eta0 = alpha_c - 1.0 / (gamma0 * gamma0)
    zc, skipped = O.apply_phase_advance(z, dE, eta0, h, beta0, E0, rf, phis)
    assert not skipped
    return zu, zc


def test_drift_zero_slip_leaves_z_on_the_phase_round_trip():
    # alpha_c = 1/gamma_p^2 exactly (gamma_p = 32): the particle's slip factor
    # is zero, so the phase does not advance and z only makes the round trip
    # z -> phi -> z.  With gamma0 in place of gamma_p the slip would be
    # 1/1024 - 1/900 and z would move by ~1e-4 of its size.
    mass, gamma0, alpha_c = 1.0, 30.0, 1.0 / 1024.0
    z = np.array([1.3e-3, -2.7e-3, 4.1e-3])
    dE = np.full(3, 2.0)                       # gamma_p = 30 + 2/1 = 32
    rf, phis, E0, beta0, h = 9.987, 2.82377, 30.0, 0.9994, 372
    zu, _ = _drift_pair(z, dE, mass, gamma0, alpha_c, h, beta0, E0, rf, phis)
    ulp_phase = np.spacing(phis) / rf
    assert np.all(np.abs(zu - z) <= 4 * ulp_phase + 4 * np.spacing(z))
    # a particle at gamma0 itself (dE = 0) drifts with the constant slip factor,
    # bit for bit
    p = W.ring_3gev_lattice()
    g0 = p["E0"] / p["mass"]
    b0 = math.sqrt(1.0 - 1.0 / (g0 * g0))
    rf3 = O.calc_rf_factor(p["freq_rf"], b0)
    z3 = np.linspace(-8e-3, 8e-3, 17)
    zu3, zc3 = _drift_pair(z3, np.zeros(17), p["mass"], g0, p["alpha_c"], p["harmonic"], b0,
                           p["E0"], rf3, p["phi_s"])
    np.testing.assert_array_equal(zu3, zc3)


def test_drift_energy_dependent_slip_follows_its_expansion():
    # eta(gamma_p) - eta(gamma0) = 1/gamma0^2 - 1/gamma_p^2
    #   = 2 x/gamma0^3 - 3 x^2/gamma0^4 + ...,  x = dE/m,
    # so the update_eta drift and the constant-eta drift differ by
    #   dz = -(2 pi h / (beta0^2 E0 rf)) (2 x/gamma0^3 - 3 x^2/gamma0^4) dE.
    # The expansion (not the oracle's expression) is the expectation; its
    # third-order remainder is ~ (x/gamma0)^2 ~ 1e-6 of the difference.
    p = W.ring_3gev_lattice()
    m, E0, h = p["mass"], p["E0"], p["harmonic"]
    g0 = E0 / m
    b0 = math.sqrt(1.0 - 1.0 / (g0 * g0))
    rf = O.calc_rf_factor(p["freq_rf"], b0)
    z = np.full(6, 1.0e-3)
    dE = np.array([-9e6, -3e6, -1e6, 1e6, 3e6, 9e6])
    zu, zc = _drift_pair(z, dE, m, g0, p["alpha_c"], h, b0, E0, rf, p["phi_s"])
    k = (2.0 * math.pi * h) / (b0 * b0 * E0 * rf)
    x = dE / m
    expect = -k * (2.0 * x / g0 ** 3 - 3.0 * x * x / g0 ** 4) * dE
    got = zu - zc
    # rounding of z itself (ulp ~ 2e-19) against differences of 1e-12 .. 1e-10
    np.testing.assert_allclose(got, expect, rtol=2e-5, atol=64 * np.spacing(1e-3))
    # the slope is even in dE at first order: eta rises with energy above
    # transition, so both signs of dE move z the same way
    assert np.all(np.sign(got) == np.sign(-k))


# --------------------------------------------------------------------------
# wake function (wakefield.jl:115-118) with the constants of evolution.jl:276-280
# --------------------------------------------------------------------------

def _pipe_wake_constants(b):
    # Z0 = 120 pi, W(0) = Z0 c / (pi b) = 120 c / b; k_p = 30 -> ws = sqrt(60 / b);
    # c tau = 4 mm -> 1/(c tau) = 250 per metre
    wf = ((120 * math.pi) * W.SPEED_LIGHT) / (math.pi * b)
    ws = math.sqrt(max(0.0, (2 * 30.0) / b))
    return wf, ws, 1.0 / 4e-3


@pytest.mark.parametrize("b", [0.25e-3, 1e-3, 0.02])
def test_wake_function_value_zero_and_decay(b):
    wf, ws, inv_ctau = _pipe_wake_constants(b)
    # W(0) = Z0 c / (pi b) = 120 c / b
    assert O.wake_function(0.0, wf, ws, inv_ctau) == pytest.approx(120 * W.SPEED_LIGHT / b,
                                                                   rel=4e-16)
    # causality: nothing ahead of the source
    for z in (1e-12, 1e-3, 0.5):
        assert O.wake_function(z, wf, ws, inv_ctau) == 0.0
    # the first zero of the cosine, z1 = -pi / (2 ws): W changes sign there
    z1 = -math.pi / (2 * ws)
    assert O.wake_function(z1 * (1 - 1e-9), wf, ws, inv_ctau) > 0
    assert O.wake_function(z1 * (1 + 1e-9), wf, ws, inv_ctau) < 0
    assert abs(O.wake_function(z1, wf, ws, inv_ctau)) <= 1e-15 * wf
    # cosine extrema z_k = -k pi / ws: consecutive maxima of one sign (k, k+2)
    # are in the ratio W(z_k) / W(z_{k+2}) = exp(2 pi / (ws c tau))
    for k in (0, 1, 3):
        zk = -k * math.pi / ws
        zk2 = -(k + 2) * math.pi / ws
        ratio = O.wake_function(zk, wf, ws, inv_ctau) / O.wake_function(zk2, wf, ws, inv_ctau)
        assert ratio == pytest.approx(math.exp(2 * math.pi * 250.0 / ws), rel=1e-12)
        # and the extremum's size is the damped amplitude wf exp(z_k / c tau)
        assert abs(O.wake_function(zk, wf, ws, inv_ctau)) == pytest.approx(
            wf * math.exp(zk * 250.0), rel=1e-13)


def test_wake_function_integral_matches_closed_form():
    # int_{-inf}^0 wf e^{z/ct} cos(ws z) dz = wf (1/ct) / ((1/ct)^2 + ws^2)
    b = 0.02
    wf, ws, a = _pipe_wake_constants(b)
    zz = np.linspace(-40.0 / a, 0.0, 2_000_001)
    w = np.array([O.wake_function(v, wf, ws, a) for v in zz[::1000]])   # coarse sanity
    assert np.all(np.isfinite(w))
    f = np.vectorize(lambda v: O.wake_function(v, wf, ws, a))
    vals = f(zz)
    integral = np.trapezoid(vals, zz)
    assert integral == pytest.approx(wf * a / (a * a + ws * ws), rel=1e-6)


# --------------------------------------------------------------------------
# delta kernel (utils.jl:58-66)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("sigma", [0.15, 1.0, 3e-3])
def test_delta_is_the_unit_normal_density(sigma):
    peak = 1.0 / (math.sqrt(2 * math.pi) * sigma)
    assert O.delta(0.0, sigma) == pytest.approx(peak, rel=2.3e-16)
    # one standard deviation out: e^{-1/2} of the peak; two: e^{-2}
    assert O.delta(sigma, sigma) / O.delta(0.0, sigma) == pytest.approx(math.exp(-0.5), rel=1e-15)
    assert O.delta(2 * sigma, sigma) / O.delta(0.0, sigma) == pytest.approx(math.exp(-2.0),
                                                                           rel=1e-15)
    assert O.delta(-0.7 * sigma, sigma) == O.delta(0.7 * sigma, sigma)
    # unit mass and variance sigma^2 (Gauss-Legendre quadrature in mpmath)
    f = lambda x: mp.mpf(O.delta(float(x), sigma))
    mass = mp.quad(f, [-12 * sigma, -4 * sigma, 0, 4 * sigma, 12 * sigma])
    var = mp.quad(lambda x: x * x * f(x), [-12 * sigma, -4 * sigma, 0, 4 * sigma, 12 * sigma])
    assert float(mass) == pytest.approx(1.0, rel=1e-13)
    assert float(var) == pytest.approx(sigma * sigma, rel=1e-13)


# --------------------------------------------------------------------------
# wake current heuristic (evolution.jl:373-386)
# --------------------------------------------------------------------------

def test_wake_current_scaling_laws():
    base = dict(n_global=123_456, E0=3e9, radius=37.24, sigma_z=4e-3, eta0=1.1e-3, sigma_E0=3e6)
    I0 = O.wake_current(**base)

    def I(**kw):
        return O.wake_current(**dict(base, **kw))
    # powers of two scale exactly in binary floating point
    assert I(sigma_z=2 * base["sigma_z"]) == 2 * I0                 # proportional to sigma_z
    assert I(E0=2 * base["E0"]) == I0 / 2                           # to 1/E0
    assert I(radius=2 * base["radius"]) == I0 / 2                   # to 1/R
    assert I(eta0=2 * base["eta0"]) == I0 / 2                       # to 1/eta0
    assert I(sigma_E0=2 * base["sigma_E0"]) == I0 / 4               # to 1/sigma_E0^2
    assert I(sigma_E0=3 * base["sigma_E0"]) == pytest.approx(I0 / 9, rel=4e-16)
    assert I(eta0=-base["eta0"]) == -I0                             # below transition: sign
    # the particle factor 1e11 N / 10^floor(log10 N) lies in [1e11, 1e12):
    # N and 10 N give the same current
    for N in (7, 123_456, 10 ** 7):
        assert I(n_global=10 * N) == pytest.approx(I(n_global=N), rel=1e-15)
    assert I(n_global=2 * 10 ** 6) == pytest.approx(2 * I(n_global=10 ** 6), rel=1e-15)


def test_wake_current_absolute_value():
    # N = 1e5 -> particle factor 1e11; I = 1e11 / E0 / (2 pi R) * sz / (eta sE^2)
    E0, R, sz, eta, sE = 3e9, 37.24, 4e-3, 1.1e-3, 3e6
    expect = mp.mpf(10) ** 11 / E0 / (2 * mp.pi * R) * sz / (eta * mp.mpf(sE) ** 2)
    assert O.wake_current(100_000, E0, R, sz, eta, sE) == pytest.approx(float(expect), rel=1e-15)
    # N = 1: log10 = 0, factor 1e11 * 1
    assert O.wake_current(1, 1.0, 1.0 / (2 * math.pi), 1.0, 1.0, 1.0) == pytest.approx(1e11,
                                                                                      rel=1e-15)
