"""Branch semantics of longitudinal_evolve on the GPU against the oracle:
the serial branch (use_mpi = false), non-finite particles, the charge
option's line density and the error paths of the C ABI.
"""
import math

import numpy as np
import pytest

from helpers import COORD_RTOL, POTENTIAL_RTOL, W, normwise, to_sim_params
from oracle import oracle as O

import haissinskijulia_amd as H
from haissinskijulia_amd._lib import HM_ERR_ARG, HaissinskiError

pytestmark = pytest.mark.gpu


def _evolve(z0, e0, nbins, p, use_mpi=True, **kw):
    part = H.Particles(z0.copy(), e0.copy())
    bufs = H.create_simulation_buffers(z0.size, max(nbins, 1))
    r = H.longitudinal_evolve(part, to_sim_params(p), bufs, H.MI355X.default(), use_mpi,
                              H.TrackingOptions(**kw))
    return part.coordinates.z, part.coordinates.ΔE, r


@pytest.mark.parametrize("flags", [(False, True, True, True, True), (False, False, True, True, False)])
def test_serial_branch_matches_oracle(flags):
    """evolution.jl:245-253,437-447,461-467: compute_std (two passes) for
    sigma_E0 (the QE amplitude) and the final spreads.  The bunch's energy
    mean is 100 sigma_E off zero, where the MPI branch's one-pass spread
    loses digits against the two-pass one (~1e-12 relative here)."""
    p = W.ring_3gev_lattice(40, flags=flags)
    n = 20_001
    z0, e0 = O.generate_gaussian(31, 2e-4, 4e-3, 3e8, 3e6, n)
    nrm = np.vstack([O.normals(32, t, 0, n) for t in range(40)])
    zr, er, rr = O.evolve(z0, e0, 0, p, normals=nrm, serial=True)
    assert rr.status == 0
    z, e, (sE, sz, E0) = _evolve(z0, e0, 0, p, use_mpi=False, normals=nrm)
    print(f"serial: normwise z={normwise(z, zr):.2e} dE={normwise(e, er):.2e} "
          f"sigma_E {sE:.15e} oracle {rr.sigma_E:.15e}", flush=True)
    assert normwise(z, zr) <= COORD_RTOL and normwise(e, er) <= COORD_RTOL
    assert E0 == pytest.approx(rr.E0, rel=1e-14)
    assert sE == pytest.approx(rr.sigma_E, rel=1e-13)
    assert sz == pytest.approx(rr.sigma_z, rel=1e-13)
    # the GPU's initial spread is the two-pass one too (it sets the excitation)
    assert H.longitudinal_evolve.last.sigma_E0 == pytest.approx(O.serial_std(e0), rel=1e-14)
    # and the MPI branch's one-pass spread is measurably further off: by more
    # than ten times the serial branch's distance to the oracle
    z1, e1, (sE1, _, _) = _evolve(z0, e0, 0, p, use_mpi=True, normals=nrm)
    assert abs(sE1 - sE) > 10 * abs(sE - rr.sigma_E) and abs(sE1 - sE) > 1e-14 * sE


def test_serial_branch_edge_cases():
    p = W.ring_3gev_lattice(3, flags=(False, True, True, True, True))
    e = np.empty(0)
    _, _, (sE, sz, E0) = _evolve(e, e, 0, p, use_mpi=False)
    assert E0 == p["E0"] and sE == 0.0 and sz == 0.0     # no E0 update of an empty bunch
    _, _, (_, _, E0m) = _evolve(e, e, 0, p, use_mpi=True)
    assert math.isnan(E0m)                              # the MPI branch divides 0 by 0
    z0, e0 = O.generate_gaussian(1, 0.0, 4e-3, 0.0, 3e6, 100)
    with pytest.raises(HaissinskiError) as ei:           # evolution.jl:267 with comm = nothing
        _evolve(z0, e0, 64, dict(p, use_wakefield=True), use_mpi=False)
    assert ei.value.status == HM_ERR_ARG and "MethodError" in str(ei.value)


def test_charge_with_a_delta_kernel_is_rejected():
    p = W.ring_3gev_lattice(2)
    z0, e0 = O.generate_gaussian(1, 0.0, 4e-3, 0.0, 3e6, 100)
    with pytest.raises(HaissinskiError) as ei:
        _evolve(z0, e0, 64, p, charge=1e-9, resonators=[W.RESONATOR_BB], delta_sigma=0.15)
    assert ei.value.status == HM_ERR_ARG


@pytest.mark.parametrize("update_E0", [False, True])
def test_non_finite_particle_matches_oracle(update_E0):
    """One particle starts with dE = NaN (QE off, so sigma_E0 enters nothing;
    the charge sets the current): its z turns NaN in the first drift, the
    histogram leaves it out, the other particles keep finite potentials, and
    with update_E0 the energy sum turns E0 NaN in the same turn as the
    oracle's.  Non-finite entries must sit in the same places."""
    p = W.ring_3gev_lattice(12, flags=(True, True, update_E0, True, False))
    n = 4001
    z0, e0 = O.generate_gaussian(5, 0.0, 4e-3, 0.0, 3e6, n)
    e0[1234] = np.nan
    kw = dict(resonators=[W.RESONATOR_BB], charge=2e-9, bin_half_width_sigmas=8.0)
    mon_o = np.zeros((12, 4))
    zr, er, rr = O.evolve(z0, e0, 256, p, seed=3, monitor=mon_o, **kw)
    assert rr.status == 0
    z, e, (sE, sz, E0) = _evolve(z0, e0, 256, p, seed=3, monitor=True, **kw)
    mon = H.longitudinal_evolve.last.monitor
    for a, b in ((z, zr), (e, er)):
        np.testing.assert_array_equal(np.isfinite(a), np.isfinite(b))
        f = np.isfinite(b)
        if f.any():
            assert normwise(a[f], b[f]) <= COORD_RTOL
    assert np.isnan(z[1234]) and np.isnan(e[1234])
    np.testing.assert_array_equal(np.isnan(mon), np.isnan(mon_o))
    np.testing.assert_array_equal(mon[:, 3], mon_o[:, 3])        # binned counts: n - 1
    assert np.all(mon[:, 3] <= n - 1)
    if update_E0:
        assert math.isnan(E0) and math.isnan(rr.E0)
    else:
        assert E0 == p["E0"] and rr.E0 == p["E0"]


def test_interpolation_kick_of_non_finite_z():
    """zero_alloc_interpolation! (wakefield.jl:64-78): clamp keeps a NaN z,
    whose weight and kick are NaN; +-Inf clamp to the end centres."""
    nb = 64
    _, centres, _ = O.bin_grid(0.02, nb)
    P = np.linspace(-3.0, 5.0, nb) ** 2
    z = np.array([np.nan, np.inf, -np.inf, 0.0013, centres[-1], 1.0])
    dE = np.zeros(z.size)
    part = H.Particles(z.copy(), dE.copy())
    H.api.interpolate_kick(part, centres, P)
    expect = -O.interpolate(z, centres, P)
    np.testing.assert_array_equal(np.isnan(part.coordinates.ΔE), np.isnan(expect))
    f = ~np.isnan(expect)
    np.testing.assert_array_equal(part.coordinates.ΔE[f], expect[f])
    assert np.isnan(part.coordinates.ΔE[0])


def test_lag_wake_potential_is_causal_on_the_gpu():
    """A resonator wake (by lag) from a clump of charge at z = +4 mm, probed by
    a few particles spread over +-12 mm: one turn without RF voltage, SR, QE
    or energy spread (the drift then moves nothing), so each probe's dE is
    minus its wake kick.  Probes behind the clump (lower z) are kicked; probes
    ahead of it feel only the probes' own small charge."""
    p = dict(W.ring_3gev_lattice(1, flags=(True, True, False, False, False)), voltage=0.0)
    rng = np.random.default_rng(4)
    probes = np.linspace(-12e-3, 12e-3, 401)
    z0 = np.concatenate([rng.normal(4e-3, 5e-4, 60_000), probes])
    e0 = np.zeros(z0.size)
    kw = dict(resonators=[W.RESONATOR_BB], charge=5e-9, bin_half_width_sigmas=20.0)
    zr, er, rr = O.evolve(z0, e0, 512, p, seed=1, **kw)
    assert rr.status == 0
    z, e, _ = _evolve(z0, e0, 512, p, seed=1, **kw)
    # dE is the kick alone here, so the potential's tolerance applies
    assert normwise(e, er) <= POTENTIAL_RTOL
    kick = -e[60_000:]
    ahead = probes > 4e-3 + 6 * 5e-4
    behind = (probes > -1e-3) & (probes < 1e-3)
    print(f"max |kick| ahead {np.abs(kick[ahead]).max():.3e}, "
          f"min |kick| behind {np.abs(kick[behind]).min():.3e}", flush=True)
    assert np.abs(kick[ahead]).max() <= 0.02 * np.abs(kick[behind]).min()


@pytest.mark.parametrize("nbins", [2, 1024, 8192])
def test_cloud_in_cell_sum_with_particles_off_the_grid(nbins):
    """With update_E0, a bunch of at most 8192 bins takes the sum of the
    interpolated potential (evolution.jl:419-429) from the pass's cloud-in-cell
    deposit instead of a second read of z.  Particles beyond the end centres
    clamp to the end cells as the kick's interpolation does
    (wakefield.jl:64-78): bins over +-1.6 sigma_z0 leave a tenth of the
    bunch outside the grid.  Two bins are the smallest grid with a cell."""
    p = W.ring_3gev_lattice(30)
    n = 20_001
    z0, e0 = O.generate_gaussian(12, 0.0, 4e-3, 0.0, 3e6, n)
    assert np.mean(np.abs(z0) > 1.6 * 4e-3) > 0.09
    kw = dict(resonators=[W.RESONATOR_BB], charge=5e-9, bin_half_width_sigmas=1.6)
    nrm = np.vstack([O.normals(13, t, 0, n) for t in range(30)])
    zr, er, rr = O.evolve(z0, e0, nbins, p, normals=nrm, **kw)
    assert rr.status == 0
    z, e, (sE, sz, E0) = _evolve(z0, e0, nbins, p, normals=nrm, **kw)
    print(f"nbins {nbins}: normwise z={normwise(z, zr):.2e} dE={normwise(e, er):.2e} "
          f"E0 {E0!r} oracle {rr.E0!r}", flush=True)
    assert normwise(z, zr) <= COORD_RTOL and normwise(e, er) <= COORD_RTOL
    assert E0 == pytest.approx(rr.E0, rel=1e-14)
    assert sE == pytest.approx(rr.sigma_E, rel=1e-10) and sz == pytest.approx(rr.sigma_z, rel=1e-10)
