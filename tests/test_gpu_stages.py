"""Stage-level parity of the HIP kernels with the oracle, through the C ABI.

Bit-exact where the stage is integer/index work or IEEE +,-,*,/ on the same
inputs (histogram, interpolation, QE, SR, drift); within a few ulp where a
transcendental of the device math library enters (RF kick's sin, the normals'
log/sincos); within 1e-10 norm-wise for the FFT convolution.
"""
import math

import numpy as np
import pytest

from helpers import POTENTIAL_RTOL, W, normwise
from oracle import oracle as O

import haissinskijulia_amd as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = H.MI355X(0)
    info = d.info()
    print("device", info, flush=True)
    assert info["arch"].startswith("gfx950")
    yield d
    d.close()


def _bunch(n, seed=3):
    return O.generate_gaussian(seed, 0.0, 4e-3, 0.0, 3e6, n)


def test_rf_kick(dev):
    p = W.ring_3gev_lattice()
    z, dE = _bunch(10_001)
    rf = O.calc_rf_factor(p["freq_rf"], 0.99999998)
    sp = math.sin(p["phi_s"])
    ref = O.rf_kick(p["voltage"], sp, rf, p["phi_s"], z, dE)
    part = H.Particles(z.copy(), dE.copy())
    H.rf_kick(p["voltage"], sp, rf, p["phi_s"], part, None, dev)
    # one ulp of sin scaled by the voltage
    assert np.max(np.abs(part.coordinates.ΔE - ref)) <= 4 * p["voltage"] * 2.0 ** -52


def test_quantum_excitation_and_sr_bit_exact(dev):
    z, dE = _bunch(4097)
    r = O.normals(5, 0, 0, dE.size)
    part = H.Particles(z, dE.copy())
    H.quantum_excitation(3e9, 37.24, 3e6, None, part, dev, normals=r)
    np.testing.assert_array_equal(part.coordinates.ΔE, O.quantum_excitation(3e9, 37.24, 3e6, r, dE))
    part = H.Particles(z, dE.copy())
    H.synchrotron_radiation(3e9, 37.24, part, None, dev)
    np.testing.assert_array_equal(part.coordinates.ΔE, O.synchrotron_radiation(3e9, 37.24, dE))


@pytest.mark.parametrize("update_eta", [True, False])
def test_drift_within_one_phase_ulp(dev, update_eta):
    # The kernel multiplies by 1/rf, 1/mass and (2 pi h)/denom where the
    # reference divides; the phase phi then rounds to the same double or a
    # neighbour, so z agrees to one ulp of phi_s / rf plus a few ulp of z.
    p = W.ring_3gev_lattice(flags=(True, update_eta, True, True, True))
    z, dE = _bunch(5000)
    E0 = p["E0"]
    g0 = E0 / p["mass"]
    b0 = math.sqrt(max(0.0, 1.0 - 1.0 / (g0 * g0)))
    eta0 = p["alpha_c"] - (1.0 / (g0 * g0))
    rf = O.calc_rf_factor(p["freq_rf"], b0)
    if update_eta:
        ref = O.drift_update_eta(z, dE, p["mass"], g0, p["alpha_c"], p["harmonic"], b0, E0, rf, p["phi_s"])
    else:
        ref, _ = O.apply_phase_advance(z, dE, eta0, p["harmonic"], b0, E0, rf, p["phi_s"])
    part = H.Particles(z.copy(), dE)
    skipped = H.drift(part, H.SimulationParameters.from_dict(p), E0, dev)
    assert not skipped
    bound = 2 * np.spacing(p["phi_s"]) / rf + 4 * np.spacing(np.abs(ref))
    assert np.all(np.abs(part.coordinates.z - ref) <= bound)


def test_drift_skip_condition(dev):
    p = W.ring_3gev_lattice(flags=(False, False, False, False, False))
    z, dE = _bunch(100)
    part = H.Particles(z.copy(), dE)
    # E0 = 0 -> apply_phase_advance! warns and skips (evolution.jl:151-154)
    assert H.drift(part, H.SimulationParameters.from_dict(p), 0.0, dev)
    np.testing.assert_array_equal(part.coordinates.z, z)


@pytest.mark.parametrize("nb", [1, 7, 1024, 16384, 70000])
def test_histogram_bit_exact(dev, nb):
    edges, centers, step = O.bin_grid(0.02, nb)
    rng = np.random.default_rng(nb)
    x = rng.normal(0, 0.006, 200_003)
    x[:min(nb + 1, 1000)] = edges[:min(nb + 1, 1000)]
    x[-5:] = [edges[-1], np.nan, np.inf, -np.inf, np.nextafter(edges[-1], 1.0)]
    ref = O.histogram(x, edges[0], edges[-1], 1.0 / step, nb)
    got = H.histogram(x, edges, dev)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("nb,fft_block", [(1, 0), (5, 0), (64, 0), (1000, 0), (1024, 0), (1024, 64),
                                          (4096, 0), (8192, 0), (8192, 1024), (16384, 0)])
def test_wake_potential_matches_oracle_fft(dev, nb, fft_block):
    edges, centers, step = O.bin_grid(0.03, nb)
    rng = np.random.default_rng(nb)
    counts = rng.integers(0, 5000, nb).astype(np.int64)
    b = 0.25e-3
    wf = ((120 * math.pi) * W.SPEED_LIGHT) / (math.pi * b)
    ws = math.sqrt((2 * 30.0) / b)
    wt = np.array([O.wake_function(c, wf, ws, 1 / 4e-3) for c in centers])
    lt = np.array([O.delta(c, 0.15) for c in centers])
    cur = 3.7e-11
    ref = O.wake_potential(counts, wt, lt, cur)
    got = H.wake_potential(counts, wt, lt, cur, dev, fft_block=fft_block)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"nb={nb} B={fft_block} normwise={err:.2e}", flush=True)
    assert err <= POTENTIAL_RTOL
    assert err <= 1e-13


def test_wake_potential_resonator_table_full_support(dev):
    nb = 4096
    edges, centers, step = O.bin_grid(0.024, nb)
    wt = np.array([O.resonator_wake(k * step, 10e3, 5e9, 1.0) for k in range(nb)])
    counts = np.random.default_rng(1).integers(0, 100, nb).astype(np.int64)
    ref = O.wake_potential(counts, wt, np.ones(nb), 1e-9)
    for B in (0, 256, 1024):
        got = H.wake_potential(counts, wt, np.ones(nb), 1e-9, dev, fft_block=B)
        assert np.linalg.norm(got - ref) <= 1e-13 * np.linalg.norm(ref)


def test_wake_potential_zero_counts(dev):
    got = H.wake_potential(np.zeros(64, dtype=np.int64), np.ones(64), np.ones(64), 1.0, dev)
    assert np.all(got == 0)


@pytest.mark.parametrize("nb", [1, 2, 50, 1024])
def test_interpolate_kick_bit_exact(dev, nb):
    _, centers, _ = O.bin_grid(0.01, nb)
    rng = np.random.default_rng(2)
    P = rng.normal(size=nb)
    z = np.concatenate([rng.uniform(-0.012, 0.012, 10_000), centers, [np.inf, -np.inf, 1.0, -1.0]])
    dE = rng.normal(size=z.size)
    ref = dE - O.interpolate(z, centers, P)
    part = H.Particles(z, dE.copy())
    H.interpolate_kick(part, centers, P, dev)
    np.testing.assert_array_equal(part.coordinates.ΔE, ref)


def test_device_normals_match_oracle_stream(dev):
    got = H.draw_normals(0xDEADBEEF12, 77, 3, 100_001, dev)
    ref = O.normals(0xDEADBEEF12, 77, 3, 100_001)
    np.testing.assert_allclose(got, ref, rtol=4e-15, atol=4e-15)


def test_device_generation_matches_oracle(dev):
    b = H.DeviceBunch(50_001, dev)
    b.generate(11, 4e-3, 3e6)
    zg, eg = b.download()
    b.close()
    z, dE = O.generate_gaussian(11, 0.0, 4e-3, 0.0, 3e6, 50_001)
    np.testing.assert_allclose(zg, z, rtol=4e-15, atol=1e-18)
    np.testing.assert_allclose(eg, dE, rtol=4e-15, atol=1e-8)


@pytest.mark.parametrize("n", [2, 3, 9_999, 10_000, 10_001, 250_001])
def test_generate_particles_matches_oracle(dev, n):
    """generate_particles (evolution.jl:42-133): sample-covariance MvNormal.
    The covariance sums are reduced in a different order on the GPU, so the
    Cholesky factor and the coordinates agree to rounding, not bitwise."""
    p, sE, sz, E0 = H.generate_particles(1e-4, -2e3, 4e-3, 3e6, n, 3e9, H.MASS_ELECTRON,
                                         2.8, 4.7e8, seed=77, comm=dev)
    z, dE, sEr, szr, L = O.generate_particles(77, 1e-4, -2e3, 4e-3, 3e6, n)
    assert E0 == 3e9
    if n >= 100:   # well-conditioned sample covariance
        assert normwise(p.coordinates.z - 1e-4, z - 1e-4) <= 1e-12
        assert normwise(p.coordinates.ΔE + 2e3, dE + 2e3) <= 1e-12
        assert sz == pytest.approx(szr, rel=1e-12) and sE == pytest.approx(sEr, rel=1e-12)
    else:          # m <= 3 samples: the eigenvalue lift is at rounding level
        assert np.all(np.isfinite(p.coordinates.z)) and np.all(np.isfinite(p.coordinates.ΔE))
        assert sz == pytest.approx(np.std(p.coordinates.z, ddof=1), rel=1e-12)


def test_generate_particles_one_and_zero(dev):
    p, sE, sz, E0 = H.generate_particles(0.0, 0.0, 4e-3, 3e6, 1, 3e9, H.MASS_ELECTRON, 2.8, 4.7e8,
                                         seed=5, comm=dev)
    z, dE = O.generate_gaussian(5, 0.0, 4e-3, 0.0, 3e6, 1)
    assert (sE, sz) == (3e6, 4e-3)
    np.testing.assert_allclose(p.coordinates.z, z, rtol=4e-15)
    p, sE, sz, E0 = H.generate_particles(0.0, 0.0, 4e-3, 3e6, 0, 3e9, H.MASS_ELECTRON, 2.8, 4.7e8,
                                         seed=5, comm=dev)
    assert len(p) == 0 and (sE, sz, E0) == (0.0, 0.0, 3e9)


def test_global_std_matches_oracle(dev):
    x = np.random.default_rng(4).normal(1.0, 2.0, 1_000_003)
    assert H.compute_global_std(x, dev) == pytest.approx(O.global_std(x), rel=1e-13)
    assert H.compute_global_std(np.array([1.0, 2.0, 3.0, 4.0]), dev) == math.sqrt(5 / 3)


def test_apply_wakefield_inplace_matches_oracle(dev):
    p = W.reference_benchmark_lattice()
    z, dE = _bunch(20_000)
    nb = 256
    zw = 7.5 * O.global_std(z)
    edges, centers, step = O.bin_grid(zw, nb)
    b = p["pipe_radius"]
    wf = ((120 * math.pi) * W.SPEED_LIGHT) / (math.pi * b)
    ws = math.sqrt((2 * 30.0) / b)
    wt = np.array([O.wake_function(c, wf, ws, 1 / 4e-3) for c in centers])
    lt = np.array([O.delta(c, 0.15) for c in centers])
    counts = O.histogram(z, edges[0], edges[-1], 1 / step, nb)
    cur = 2.5e-11
    P = O.wake_potential(counts, wt, lt, cur)
    ref = dE - O.interpolate(z, centers, P)
    part = H.Particles(z, dE.copy())
    bufs = H.create_simulation_buffers(z.size, nb)
    H.apply_wakefield_inplace(part, bufs, wf, ws, 4e-3, cur, 0.0, edges, centers, dev)
    assert np.max(np.abs(part.coordinates.ΔE - ref)) <= 1e-12 * np.max(np.abs(P))
