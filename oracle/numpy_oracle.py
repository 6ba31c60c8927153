"""Independent numpy restatement of the reference turn loop, for tiny inputs.

TEST INFRASTRUCTURE ONLY (a second checker written apart from the C oracle,
so that a slip in one of them shows up as a disagreement).

Follows src/evolution.jl:206-471 (MPI branch, one rank), with the stage
functions of src/rf_kick.jl:84-87, src/quantum_excitation.jl:58-71,
src/synchrotron_radiation.jl:44-56, src/wakefield.jl:22-118,166-345 and
src/utils.jl:58-66,351-401,569-606.  Bin grids are computed with exact
rational arithmetic (fractions), the convolution with numpy.fft at the
reference's padded length.
"""
from __future__ import annotations

from fractions import Fraction
import math

import numpy as np

C_LIGHT = 299792458.0
TWO_PI = 2 * math.pi


def jl_max(a, b):
    if math.isnan(a):
        return a
    if math.isnan(b):
        return b
    return b if b > a else a


def gstd(x: np.ndarray) -> float:
    """utils.jl:569-606 (one rank): one-pass sums."""
    n = x.size
    s = 0.0
    s2 = 0.0
    for v in x.tolist():
        s += v
        s2 += v * v
    if n <= 1:
        return 0.0
    mean = s / n
    var = (s2 / n) - mean * mean
    return math.sqrt(jl_max(0.0, var * n / (n - 1)))


def exact_grid(zw: float, nb: int):
    """Correctly rounded edges/centres of range(-zw, zw, length=nb+1)."""
    fz = Fraction(zw)
    edges = np.array([float(fz * (2 * k - nb) / nb) for k in range(nb + 1)])
    centers = np.array([float(fz * (2 * i + 1 - nb) / nb) for i in range(nb)])
    step = float(Fraction(2) * fz / nb)
    return edges, centers, step


def rf_kick(V, sphis, rf, phis, z, dE):
    phi = ((-z) * rf) + phis
    return dE + V * (np.sin(phi) - sphis)


def dUdE(E0, radius):
    x = E0 / 1e9
    return ((4 * 8.85e-5) * ((x * x) * x)) / radius


def quantum_excitation(E0, radius, sE0, r, dE):
    d = dUdE(E0, radius)
    exc = math.sqrt(1 - (1 - d) * (1 - d)) * sE0
    return dE + exc * r


def synchrotron_radiation(E0, radius, dE):
    return dE * (1 - dUdE(E0, radius))


def drift_update_eta(z, dE, mass, gamma0, alpha_c, h, beta0, E0, rf, phis):
    dg = dE / mass
    gp = gamma0 + dg
    eta = alpha_c - 1.0 / (gp * gp)
    coeff = ((TWO_PI * h) * eta) / ((beta0 * beta0) * E0)
    phi = -((z * rf) - phis)
    phi = phi + coeff * dE
    return ((-phi) + phis) / rf


def histogram(x, e_first, e_last, inv_step, nb):
    counts = np.zeros(nb, dtype=np.int64)
    for v in x.tolist():
        if v >= e_first and v <= e_last:
            b = int(math.floor((v - e_first) * inv_step))
            counts[min(max(b, 0), nb - 1)] += 1
    return counts


def wake_W(z, wf, ws, ictau):
    return 0.0 if z > 0 else (wf * math.exp(z * ictau)) * math.cos(ws * z)


def delta(x, s):
    return (1 / (math.sqrt(2 * math.pi) * s)) * math.exp((x * x) * (-0.5 / (s * s)))


def potential_fft(counts, wtab, ltab, current):
    nb = counts.size
    L = 1 << max(0, (2 * nb - 1).bit_length())
    nh = int(counts.sum())
    if nh == 0:
        return np.zeros(nb)
    inv_n = 1 / nh
    fw = np.zeros(L, dtype=complex)
    fl = np.zeros(L, dtype=complex)
    fw[:nb] = wtab
    fl[:nb] = ltab * (counts * inv_n)
    c = np.fft.ifft((np.fft.fft(fw) * np.fft.fft(fl)) * current)
    return c.real[:nb].copy()


def interpolate(z, centers, P):
    nb = centers.size
    if nb <= 1:
        return np.full(z.size, P[0] if nb == 1 else 0.0)
    c0, c1 = centers[0], centers[-1]
    inv = 1.0 / ((c1 - c0) / (nb - 1))
    out = np.empty(z.size)
    for i, pos in enumerate(z.tolist()):
        cl = c1 if pos > c1 else (c0 if pos < c0 else pos)
        u = (cl - c0) * inv
        fu = math.floor(u)
        k = min(max(int(fu), 0), nb - 2)
        w = u - fu
        out[i] = P[k] * (1.0 - w) + P[k + 1] * w
    return out


def wake_current(N, E0, radius, sz, eta0, sE0):
    lg = math.log10(N)
    fl = math.floor(lg) if lg >= 0 else math.ceil(lg - 1)
    pw = max(-10, min(10, fl))
    pf = (1e11 / (10.0 ** pw)) * N
    return (((pf / E0) / ((2 * math.pi) * radius)) * sz) / (eta0 * (sE0 * sE0))


def std2(x: np.ndarray) -> float:
    """compute_std (utils.jl:531-544): two passes, n - 1 in the variance."""
    n = x.size
    if n <= 1:
        return 0.0
    s = 0.0
    for v in x.tolist():
        s += v
    mu = s / n
    acc = 0.0
    for v in x.tolist():
        d = v - mu
        acc += d * d
    return math.sqrt(jl_max(0.0, acc / (n - 1)))


def evolve(z, dE, nbins, p: dict, normals, serial: bool = False):
    """All five flags as given in p; normals[t, i] are the injected QE draws.
    ``serial``: the use_mpi = false branch (no wake), two-pass spreads and no
    E0 update of an empty bunch (evolution.jl:245-253,437-447,461-467)."""
    z = np.array(z, dtype=np.float64)
    dE = np.array(dE, dtype=np.float64)
    n = z.size
    E0 = p["E0"]
    spread = std2 if serial else gstd
    sE0 = spread(dE)
    sz0 = spread(z)
    if p["use_wakefield"]:
        zw = 7.5 * sz0
        if zw < 1e-9:
            zw = 1e-6
        edges, centers, step = exact_grid(zw, nbins)
        inv_step = 1.0 / step
        wf = ((120 * math.pi) * C_LIGHT) / (math.pi * p["pipe_radius"])
        ws = math.sqrt(jl_max(0.0, (2 * 30.0) / p["pipe_radius"]))
        ictau = 1 / 4e-3
        wtab = np.array([wake_W(c, wf, ws, ictau) for c in centers])
        ltab = np.array([delta(c, 0.15) for c in centers])
    sphis = math.sin(p["phi_s"])
    for t in range(p["n_turns"]):
        g0 = E0 / p["mass"]
        b0 = math.sqrt(jl_max(0.0, 1.0 - 1.0 / (g0 * g0)))
        eta0 = p["alpha_c"] - (1.0 / (g0 * g0))
        rf = (p["freq_rf"] * TWO_PI) / (b0 * C_LIGHT)
        dE_init = dE.copy()
        dE = rf_kick(p["voltage"], sphis, rf, p["phi_s"], z, dE)
        if p["use_excitation"]:
            dE = quantum_excitation(E0, p["radius"], sE0, normals[t], dE)
        if p["SR_damping"]:
            dE = synchrotron_radiation(E0, p["radius"], dE)
        if p["update_eta"]:
            z = drift_update_eta(z, dE, p["mass"], g0, p["alpha_c"], p["harmonic"], b0, E0, rf, p["phi_s"])
        else:
            coeff = ((TWO_PI * p["harmonic"]) * eta0) / ((b0 * b0) * E0)
            phi = -((z * rf) - p["phi_s"])
            phi = phi + coeff * dE
            z = ((-phi) + p["phi_s"]) / rf
        if p["use_wakefield"]:
            sz = gstd(z)
            cur = wake_current(n, E0, p["radius"], sz, eta0, sE0)
            counts = histogram(z, edges[0], edges[-1], inv_step, nbins)
            P = potential_fft(counts, wtab, ltab, cur)
            dE = dE - interpolate(z, centers, P)
        if p["update_E0"] and not (serial and n == 0):
            s = 0.0
            for v in (dE - dE_init).tolist():
                s += v
            mean = s / n
            E0 = E0 + mean
            dE = dE - mean
    return z, dE, spread(dE), spread(z), E0
