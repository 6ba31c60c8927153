"""ctypes front end of the CPU oracle (``liboracle.so``).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import this
module.  The product package never does.

The oracle restates HaissinskiJulia's MPI-branch turn loop
(src/evolution.jl:206-471) in strict IEEE binary64; see haissinski_oracle.c
for the per-function citations and the tests that pin each function.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int64)
_up = C.POINTER(C.c_uint32)


class OrParams(C.Structure):
    _fields_ = [
        ("E0", C.c_double), ("mass", C.c_double), ("voltage", C.c_double),
        ("harmonic", C.c_int64), ("radius", C.c_double), ("pipe_radius", C.c_double),
        ("alpha_c", C.c_double), ("phi_s", C.c_double), ("freq_rf", C.c_double),
        ("n_turns", C.c_int64), ("use_wakefield", C.c_int32), ("update_eta", C.c_int32),
        ("update_E0", C.c_int32), ("SR_damping", C.c_int32), ("use_excitation", C.c_int32),
    ]


class OrOptions(C.Structure):
    _fields_ = [
        ("normals", _dp), ("seed", C.c_uint64), ("turn_offset", C.c_int64),
        ("nranks", C.c_int32), ("pad_", C.c_int32), ("wake_table", _dp),
        ("delta_sigma", C.c_double), ("bin_half_width_sigmas", C.c_double),
        ("charge", C.c_double), ("n_resonators", C.c_int32), ("serial", C.c_int32),
        ("resonators", _dp), ("monitor", _dp),
    ]


class OrResult(C.Structure):
    _fields_ = [("sigma_E", C.c_double), ("sigma_z", C.c_double), ("E0", C.c_double),
                ("turns_done", C.c_int64), ("drift_skipped", C.c_int32)]


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.or_next_power_of_two.restype = C.c_int64
        L.or_next_power_of_two.argtypes = [C.c_int64]
        L.or_bin_grid.argtypes = [C.c_double, C.c_int64, _dp, _dp, _dp]
        L.or_global_std.restype = C.c_double
        L.or_global_std.argtypes = [_dp, C.c_int64]
        L.or_serial_std.restype = C.c_double
        L.or_serial_std.argtypes = [_dp, C.c_int64]
        L.or_calc_rf_factor.restype = C.c_double
        L.or_calc_rf_factor.argtypes = [C.c_double, C.c_double]
        L.or_rf_kick.argtypes = [C.c_double] * 4 + [_dp, _dp, C.c_int64]
        L.or_quantum_excitation.argtypes = [C.c_double] * 3 + [_dp, _dp, C.c_int64]
        L.or_synchrotron_radiation.argtypes = [C.c_double, C.c_double, _dp, C.c_int64]
        L.or_drift_update_eta.argtypes = [_dp, _dp, C.c_int64, C.c_double, C.c_double, C.c_double,
                                          C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double]
        L.or_apply_phase_advance.restype = C.c_int
        L.or_apply_phase_advance.argtypes = [_dp, _dp, C.c_int64, C.c_double, C.c_int64,
                                             C.c_double, C.c_double, C.c_double, C.c_double]
        L.or_histogram.argtypes = [_dp, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_int64, _ip]
        L.or_wake_function.restype = C.c_double
        L.or_wake_function.argtypes = [C.c_double] * 4
        L.or_delta.restype = C.c_double
        L.or_delta.argtypes = [C.c_double, C.c_double]
        L.or_resonator_wake.restype = C.c_double
        L.or_resonator_wake.argtypes = [C.c_double] * 4
        L.or_fft.argtypes = [_dp, _dp, C.c_int64, C.c_int]
        L.or_wake_potential.argtypes = [_ip, C.c_int64, _dp, _dp, C.c_double, _dp]
        L.or_wake_potential_direct.argtypes = [_ip, C.c_int64, _dp, _dp, C.c_double, _dp]
        L.or_wake_potential_lag.argtypes = [_ip, C.c_int64, _dp, _dp, C.c_double, _dp]
        L.or_wake_current.restype = C.c_double
        L.or_wake_current.argtypes = [C.c_int64] + [C.c_double] * 5
        L.or_interpolate.argtypes = [_dp, C.c_int64, _dp, C.c_int64, _dp, _dp]
        L.or_philox4x32_10.argtypes = [_up, _up, _up]
        L.or_normals.argtypes = [C.c_uint64, C.c_int64, C.c_int64, C.c_int64, _dp]
        L.or_generate_gaussian.argtypes = [C.c_uint64] + [C.c_double] * 4 + [C.c_int64, _dp, _dp]
        L.or_generate_particles.argtypes = [C.c_uint64] + [C.c_double] * 4 + \
            [C.c_int64, _dp, _dp, _dp, _dp, _dp]
        L.or_evolve.restype = C.c_int
        L.or_evolve.argtypes = [_dp, _dp, C.c_int64, C.c_int64, C.POINTER(OrParams),
                                C.POINTER(OrOptions), C.POINTER(OrResult), _dp]
        _lib = L
    return _lib


def _d(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def f64(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


# ---------------------------------------------------------------------------
# Stage functions (numpy in, numpy out; inputs are not modified)
# ---------------------------------------------------------------------------

def next_power_of_two(n: int) -> int:
    return int(lib().or_next_power_of_two(n))


def bin_grid(zw: float, nb: int):
    edges = np.empty(nb + 1)
    centers = np.empty(nb)
    step = C.c_double()
    lib().or_bin_grid(zw, nb, _d(edges), _d(centers), C.byref(step))
    return edges, centers, step.value


def global_std(x) -> float:
    x = f64(x)
    return float(lib().or_global_std(_d(x), x.size))


def serial_std(x) -> float:
    x = f64(x)
    return float(lib().or_serial_std(_d(x), x.size))


def calc_rf_factor(freq_rf: float, beta: float) -> float:
    return float(lib().or_calc_rf_factor(freq_rf, beta))


def rf_kick(voltage, sin_phis, rf, phis, z, dE):
    z = f64(z); dE = f64(dE).copy()
    lib().or_rf_kick(voltage, sin_phis, rf, phis, _d(z), _d(dE), z.size)
    return dE


def quantum_excitation(E0, radius, sigma_E0, r, dE):
    r = f64(r); dE = f64(dE).copy()
    lib().or_quantum_excitation(E0, radius, sigma_E0, _d(r), _d(dE), dE.size)
    return dE


def synchrotron_radiation(E0, radius, dE):
    dE = f64(dE).copy()
    lib().or_synchrotron_radiation(E0, radius, _d(dE), dE.size)
    return dE


def drift_update_eta(z, dE, mass, gamma0, alpha_c, harmonic, beta0, E0, rf, phis):
    z = f64(z).copy(); dE = f64(dE)
    lib().or_drift_update_eta(_d(z), _d(dE), z.size, mass, gamma0, alpha_c, harmonic, beta0, E0, rf, phis)
    return z


def apply_phase_advance(z, dE, eta0, harmonic, beta0, E0, rf, phis):
    z = f64(z).copy(); dE = f64(dE)
    skipped = lib().or_apply_phase_advance(_d(z), _d(dE), z.size, eta0, harmonic, beta0, E0, rf, phis)
    return z, bool(skipped)


def histogram(x, e_first, e_last, inv_step, nb):
    x = f64(x)
    counts = np.zeros(nb, dtype=np.int64)
    lib().or_histogram(_d(x), x.size, e_first, e_last, inv_step, nb, _i(counts))
    return counts


def wake_function(z, wf, ws, inv_ctau) -> float:
    return float(lib().or_wake_function(z, wf, ws, inv_ctau))


def delta(x, sigma) -> float:
    return float(lib().or_delta(x, sigma))


def resonator_wake(s, Rs, fr, Q) -> float:
    return float(lib().or_resonator_wake(s, Rs, fr, Q))


def fft(re, im, inverse=False):
    re = f64(re).copy(); im = f64(im).copy()
    lib().or_fft(_d(re), _d(im), re.size, 1 if inverse else 0)
    return re, im


def wake_potential(counts, wake_tab, lam_tab, current, direct=False, lag=False):
    """Index convolution P_k = I sum_j W_j g_{k-j} (the reference's, wakefield.jl:231-319),
    or with ``lag`` the causal one of lag-indexed tables, P_k = I sum_j W_j g_{k+j}."""
    counts = np.ascontiguousarray(counts, dtype=np.int64)
    wake_tab = f64(wake_tab); lam_tab = f64(lam_tab)
    P = np.empty(counts.size)
    if lag:
        fn = lib().or_wake_potential_lag
    else:
        fn = lib().or_wake_potential_direct if direct else lib().or_wake_potential
    fn(_i(counts), counts.size, _d(wake_tab), _d(lam_tab), current, _d(P))
    return P


def wake_current(n_global, E0, radius, sigma_z, eta0, sigma_E0) -> float:
    return float(lib().or_wake_current(n_global, E0, radius, sigma_z, eta0, sigma_E0))


def interpolate(z, centers, P):
    z = f64(z); centers = f64(centers); P = f64(P)
    out = np.empty(z.size)
    lib().or_interpolate(_d(z), z.size, _d(centers), centers.size, _d(P), _d(out))
    return out


def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().or_philox4x32_10(c, k, o)
    return [int(v) for v in o]


def normals(seed: int, turn: int, i0: int, n: int):
    out = np.empty(n)
    lib().or_normals(seed, turn, i0, n, _d(out))
    return out


def generate_gaussian(seed, mu_z, sigma_z, mu_E, sigma_E, n):
    z = np.empty(n); dE = np.empty(n)
    lib().or_generate_gaussian(seed, mu_z, sigma_z, mu_E, sigma_E, n, _d(z), _d(dE))
    return z, dE


def generate_particles(seed, mu_z, mu_E, sigma_z, sigma_E, n):
    """generate_particles (evolution.jl:42-133) on the Philox streams:
    returns (z, dE, sigma_E, sigma_z, L) with L = (l11, l21, l22)."""
    m = max(n, 1)
    z = np.empty(m); dE = np.empty(m)
    sE = np.zeros(1); sz = np.zeros(1); L = np.zeros(3)
    lib().or_generate_particles(seed, mu_z, mu_E, sigma_z, sigma_E, n, _d(z), _d(dE), _d(sE),
                                _d(sz), _d(L))
    k = max(n, 0)
    return z[:k], dE[:k], float(sE[0]), float(sz[0]), L


# ---------------------------------------------------------------------------
# Full loop
# ---------------------------------------------------------------------------

@dataclass
class EvolveResult:
    sigma_E: float
    sigma_z: float
    E0: float
    turns_done: int
    drift_skipped: int
    status: int
    P_last: np.ndarray | None


def evolve(z, dE, nbins, params: dict, normals=None, seed=0, turn_offset=0, nranks=1,
           wake_table=None, delta_sigma=0.0, bin_half_width_sigmas=0.0, charge=0.0,
           resonators=None, monitor=None, serial=False):
    """Run the restated turn loop in place on copies; returns (z, dE, EvolveResult)."""
    z = f64(z).copy(); dE = f64(dE).copy()
    p = OrParams(**{k: params[k] for k, _ in OrParams._fields_})
    keep = []
    o = OrOptions()
    if normals is not None:
        nrm = f64(normals)
        assert nrm.size >= params["n_turns"] * z.size
        keep.append(nrm)
        o.normals = _d(nrm)
    o.seed = seed
    o.turn_offset = turn_offset
    o.nranks = nranks
    if wake_table is not None:
        wt = f64(wake_table)
        keep.append(wt)
        o.wake_table = _d(wt)
    o.delta_sigma = delta_sigma
    o.bin_half_width_sigmas = bin_half_width_sigmas
    o.charge = charge
    if resonators:
        rs = f64(np.asarray(resonators, dtype=np.float64).reshape(-1))
        keep.append(rs)
        o.n_resonators = rs.size // 3
        o.resonators = _d(rs)
    o.serial = int(bool(serial))
    if monitor is not None:   # float64 [n_turns, 4], filled in place
        assert monitor.dtype == np.float64 and monitor.size >= 4 * params["n_turns"]
        o.monitor = _d(monitor)
    res = OrResult()
    P = np.zeros(max(nbins, 1))
    st = lib().or_evolve(_d(z), _d(dE), z.size, nbins, C.byref(p), C.byref(o), C.byref(res), _d(P))
    return z, dE, EvolveResult(res.sigma_E, res.sigma_z, res.E0, res.turns_done,
                               res.drift_skipped, st, P[:nbins] if params["use_wakefield"] else None)
