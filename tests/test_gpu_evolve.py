"""Whole-loop parity of longitudinal_evolve on the GPU with the oracle.

Level 1 of the parity plan: identical initial coordinates and identical
injected QE normals; coordinates after 1 and 100 turns agree to 1e-12
(max |difference| / rms of the oracle's coordinate).  The device Philox path
is checked the same way against the oracle's restatement of that stream.
"""
import math

import numpy as np
import pytest

from helpers import COORD_RTOL, W, normwise, to_sim_params
from oracle import oracle as O

import haissinskijulia_amd as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    return H.MI355X.default()


def _run_both(p, n, nbins, seed=1, normals=True, opts=None, oracle_kw=None, turn_offset=0):
    z0, e0 = O.generate_gaussian(seed, 0.0, 4e-3, 0.0, 3e6, n)
    nrm = None
    if normals:
        nrm = np.vstack([O.normals(seed + 100, t, 0, n) for t in range(p["n_turns"])]) \
            if p["n_turns"] else np.zeros((0, n))
    okw = dict(oracle_kw or {})
    zr, er, rr = O.evolve(z0, e0, nbins, p, normals=nrm, seed=seed + 100, turn_offset=turn_offset, **okw)
    assert rr.status == 0
    part = H.Particles(z0.copy(), e0.copy())
    bufs = H.create_simulation_buffers(n, max(nbins, 1))
    o = opts or H.TrackingOptions()
    o.seed = seed + 100
    o.turn_offset = turn_offset
    if normals:
        o.normals = nrm
    sE, sz, E0 = H.longitudinal_evolve(part, to_sim_params(p), bufs, H.MI355X.default(), True, o)
    return (zr, er, rr), (part.coordinates.z, part.coordinates.ΔE, sE, sz, E0)


def _check(ref, got, rtol=COORD_RTOL):
    (zr, er, rr), (zg, eg, sE, sz, E0) = ref, got
    ez, ee = normwise(zg, zr), normwise(eg, er)
    print(f"normwise z={ez:.2e} dE={ee:.2e}  E0 rel={abs(E0 - rr.E0) / abs(rr.E0):.2e}", flush=True)
    assert ez <= rtol and ee <= rtol
    assert E0 == pytest.approx(rr.E0, rel=1e-14)
    assert sE == pytest.approx(rr.sigma_E, rel=1e-10)
    assert sz == pytest.approx(rr.sigma_z, rel=1e-10)


FLAGS = [(True, True, True, True, True), (True, False, True, True, True),
         (True, True, False, True, True), (False, True, True, True, True),
         (True, True, True, False, False), (False, False, False, False, False)]


@pytest.mark.parametrize("turns", [1, 100])
@pytest.mark.parametrize("flags", FLAGS)
def test_reference_lattice_injected_normals(dev, turns, flags):
    p = W.reference_benchmark_lattice(turns, flags)
    print("case", turns, flags, flush=True)
    _check(*_run_both(p, 4096, 256))


@pytest.mark.parametrize("n,nbins", [(4097, 100), (1000, 1024), (3, 7), (2, 1), (50_000, 4096)])
def test_ring_lattice_shapes(dev, n, nbins):
    p = W.ring_3gev_lattice(n_turns=100)
    print("case", n, nbins, flush=True)
    _check(*_run_both(p, n, nbins))


def test_device_philox_stream_matches_oracle(dev):
    p = W.ring_3gev_lattice(n_turns=100)
    _check(*_run_both(p, 8192, 512, normals=False, turn_offset=12345))


def test_turn_monitor_matches_oracle(dev):
    # the turn-by-turn record (hm_options.monitor): centroid and sigma_z after
    # the drift, E0 after the turn, binned count; injected normals, 40 turns
    turns, n, nb = 40, 20_000, 512
    p = W.ring_3gev_lattice(n_turns=turns)
    res = [W.RESONATOR_BB]
    z0, e0 = O.generate_gaussian(3, 0.0, 4e-3, 0.0, 3e6, n)
    nrm = np.vstack([O.normals(103, t, 0, n) for t in range(turns)])
    mon_o = np.zeros((turns, 4))
    O.evolve(z0, e0, nb, p, normals=nrm, seed=103, resonators=res, charge=5e-9, monitor=mon_o)
    b = H.DeviceBunch(n, dev)
    b.upload(z0, e0)
    r = b.evolve(H.SimulationParameters.from_dict(p), nb,
                 H.TrackingOptions(seed=103, normals=nrm, resonators=res, charge=5e-9,
                                   monitor=True))
    b.close()
    mon = r.monitor
    print("case monitor", turns, n, nb, flush=True)
    assert mon.shape == (turns, 4)
    assert np.array_equal(mon[:, 3], mon_o[:, 3])          # binned counts: exact
    assert np.allclose(mon[:, 2], mon_o[:, 2], rtol=1e-14, atol=0.0)   # E0
    sz = np.abs(mon_o[:, 1])
    assert np.all(np.abs(mon[:, 0] - mon_o[:, 0]) <= 1e-12 * sz)      # centroid, per sigma_z
    assert np.allclose(mon[:, 1], mon_o[:, 1], rtol=1e-12, atol=0.0)  # sigma_z


def test_large_bunch_matches_oracle(dev):
    # 1.7e7 particles: tiles of 9216 pairs (nine per thread), so the
    # interpolation sums run their eight-deep ring with a partial second group,
    # and the per-bunch reductions span 15 slices of tile sums; at 16384 bins
    # the interpolation window takes more than half the LDS (one workgroup per
    # CU) and the pass's window is partial (gathers issued first).
    p = W.ring_3gev_lattice(n_turns=2)
    res = [W.RESONATOR_BB]
    o = H.TrackingOptions(resonators=res, charge=5e-9)
    okw = dict(resonators=res, charge=5e-9)
    print("case 1.7e7 particles, 16384 bins, 2 turns", flush=True)
    _check(*_run_both(p, 17_000_000, 16384, normals=False, opts=o, oracle_kw=okw))


@pytest.mark.parametrize("nbins,normals,turns", [(16384, False, 30), (16384, True, 30),
                                                 (65536, False, 10)])
def test_partial_window_kicks_match_oracle(dev, nbins, normals, turns):
    # 16384 or 65536 bins over +-4 sigma_z0: the pass's LDS window of P (12160
    # of 16384 entries, +-3 sigma_z0; or 20352 of 65536 with the histogram in
    # global memory, +-1.2 sigma_z0) leaves the bunch's tails outside it, and
    # those particles gather P from global memory.  With the device normals
    # at 16384 bins the all-flags pass issues the gathers first; with injected
    # normals, and at 65536 bins, the generic pass prefetches first.  At 65536
    # bins a bin holds a third of a particle; the potential's bin-to-bin noise
    # amplifies rounding differences faster, so that case runs 10 turns.
    p = W.ring_3gev_lattice(n_turns=turns)
    res = [W.RESONATOR_BB]
    o = H.TrackingOptions(resonators=res, charge=5e-9, bin_half_width_sigmas=4.0)
    okw = dict(resonators=res, charge=5e-9, bin_half_width_sigmas=4.0)
    print("case", nbins, "normals" if normals else "philox", flush=True)
    _check(*_run_both(p, 20_000, nbins, normals=normals, opts=o, oracle_kw=okw))


@pytest.mark.parametrize("nbins", [8192, 16384])
def test_pass_variants_agree_bitwise(dev, nbins):
    # The all-flags pass (device Philox normals) against the generic pass fed
    # the device's own normal stream (draw_normals, injected): the same
    # arithmetic, so the same bits.  Bins over +-2 sigma_z0 leave one particle
    # in seven outside the LDS window (+-1.5 sigma_z0) at 16384 bins: gather
    # first against prefetch first.  At 8192 bins the pass also deposits the
    # interpolation weights (cloud in cell), whose counters leave room for a
    # window of 8062 of the 8192 entries: the same two orders of loads.
    n, turns, seed = 20_000, 30, 101
    p = to_sim_params(W.ring_3gev_lattice(n_turns=turns))
    z0, e0 = O.generate_gaussian(1, 0.0, 4e-3, 0.0, 3e6, n)
    kw = dict(resonators=[W.RESONATOR_BB], charge=5e-9, bin_half_width_sigmas=2.0, seed=seed)
    out = []
    for inject in (False, True):
        o = H.TrackingOptions(**kw)
        if inject:
            o.normals = np.vstack([H.draw_normals(seed, t, 0, n, dev) for t in range(turns)])
        part = H.Particles(z0.copy(), e0.copy())
        r = H.longitudinal_evolve(part, p, H.create_simulation_buffers(n, nbins), dev, True, o)
        out.append((part.coordinates.z.copy(), part.coordinates.ΔE.copy(), r))
    (za, ea, ra), (zb, eb, rb) = out
    assert np.array_equal(za, zb) and np.array_equal(ea, eb)
    assert tuple(ra) == tuple(rb)


def test_extensions_resonators_charge_half_width(dev):
    p = W.ring_3gev_lattice(n_turns=100)
    res = [W.RESONATOR_BB, W.RESONATOR_LOWQ]
    opts = H.TrackingOptions(resonators=res, charge=1e-9, bin_half_width_sigmas=6.0, delta_sigma=-1.0)
    ref, got = _run_both(p, 10_000, 1024, opts=opts,
                         oracle_kw=dict(resonators=res, charge=1e-9, bin_half_width_sigmas=6.0,
                                        delta_sigma=-1.0))
    _check(ref, got)
    # the wake changed the outcome (it is not a no-op)
    p_nowake = dict(p, use_wakefield=False)
    ref2, _ = _run_both(p_nowake, 10_000, 1024)
    assert normwise(ref[0][1], ref2[0][1]) > 1e-6


@pytest.mark.parametrize("fft_block", [64, 256])
def test_fft_block_choice_does_not_change_result(dev, fft_block):
    p = W.ring_3gev_lattice(n_turns=50)
    opts = H.TrackingOptions(fft_block=fft_block)
    _check(*_run_both(p, 20_000, 2048, opts=opts))


def test_runs_are_bitwise_reproducible(dev):
    p = H.SimulationParameters.from_dict(W.ring_3gev_lattice(n_turns=200))
    z0, e0 = O.generate_gaussian(9, 0.0, 4e-3, 0.0, 3e6, 100_000)
    outs = []
    for _ in range(2):
        part = H.Particles(z0.copy(), e0.copy())
        r = H.longitudinal_evolve(part, p, H.create_simulation_buffers(z0.size, 1024), dev, True,
                                  H.TrackingOptions(seed=5))
        outs.append((part.coordinates.z.copy(), part.coordinates.ΔE.copy(), r))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])
    assert outs[0][2] == outs[1][2]


@pytest.mark.parametrize("turns", [15, 16, 50, 161])
def test_graph_replays_equal_direct_launches(dev, turns):
    """Without timing, blocks of 16 turns run as replays of one captured HIP
    graph; with timing every turn is launched directly.  Same bits."""
    p = H.SimulationParameters.from_dict(W.ring_3gev_lattice(n_turns=turns))
    z0, e0 = O.generate_gaussian(21, 0.0, 4e-3, 0.0, 3e6, 30_001)
    outs = []
    for timing in (True, False, False):
        b = H.DeviceBunch(z0.size, dev)
        b.upload(z0, e0)
        r = b.evolve(p, 1024, H.TrackingOptions(seed=8, resonators=[W.RESONATOR_BB], charge=2e-9,
                                                timing=timing))
        outs.append((*b.download(), r))
        b.close()
    for z, e, r in outs[1:]:
        np.testing.assert_array_equal(z, outs[0][0])
        np.testing.assert_array_equal(e, outs[0][1])
        assert (r.sigma_E, r.sigma_z, r.E0, r.turns_done) == \
               (outs[0][2].sigma_E, outs[0][2].sigma_z, outs[0][2].E0, turns)


def test_zero_turns_returns_initial_spreads(dev):
    p = W.ring_3gev_lattice(n_turns=0)
    z0, e0 = O.generate_gaussian(2, 0.0, 4e-3, 0.0, 3e6, 1000)
    part = H.Particles(z0.copy(), e0.copy())
    sE, sz, E0 = H.longitudinal_evolve(part, to_sim_params(p), H.create_simulation_buffers(1000, 64), dev)
    assert E0 == p["E0"]
    assert sE == pytest.approx(O.global_std(e0), rel=1e-14)
    assert sz == pytest.approx(O.global_std(z0), rel=1e-14)
    np.testing.assert_array_equal(part.coordinates.z, z0)


def test_error_codes_match_reference_errors(dev):
    p = W.ring_3gev_lattice(n_turns=1)
    z0, e0 = O.generate_gaussian(2, 0.0, 4e-3, 0.0, 3e6, 100)
    part = H.Particles(z0.copy(), e0.copy())
    with pytest.raises(H.HaissinskiError) as ei:
        H.longitudinal_evolve(part, to_sim_params(dict(p, pipe_radius=0.0)),
                              H.create_simulation_buffers(100, 16), dev)
    assert ei.value.status == -4 and "pipe_radius must be positive" in str(ei.value)
    with pytest.raises(H.HaissinskiError) as ei:
        H.longitudinal_evolve(H.Particles(np.empty(0), np.empty(0)), to_sim_params(p),
                              H.create_simulation_buffers(0, 16), dev)
    assert ei.value.status == -2
    r = H.longitudinal_evolve(H.Particles(np.empty(0), np.empty(0)),
                              to_sim_params(dict(p, use_wakefield=False)),
                              H.create_simulation_buffers(0, 16), dev)
    assert math.isnan(r[2])


def test_large_bunch_matches_oracle_for_three_turns(dev):
    # full-size parity (1e7 particles, 8192 bins), oracle on 16 threads
    p = W.ring_3gev_lattice(n_turns=3)
    n = 10_000_000
    z0, e0 = O.generate_gaussian(7, 0.0, 4e-3, 0.0, 3e6, n)
    res = [W.RESONATOR_BB]
    okw = dict(resonators=res, charge=5e-9, bin_half_width_sigmas=8.0, delta_sigma=-1.0)
    zr, er, rr = O.evolve(z0, e0, 8192, p, seed=7, nranks=16, **okw)
    part = H.Particles(z0, e0.copy())
    opts = H.TrackingOptions(seed=7, resonators=res, charge=5e-9, bin_half_width_sigmas=8.0,
                             delta_sigma=-1.0)
    sE, sz, E0 = H.longitudinal_evolve(part, to_sim_params(p), H.create_simulation_buffers(n, 8192),
                                       dev, True, opts)
    ez, ee = normwise(part.coordinates.z, zr), normwise(part.coordinates.ΔE, er)
    print(f"1e7: normwise z={ez:.2e} dE={ee:.2e}", flush=True)
    assert ez <= COORD_RTOL and ee <= COORD_RTOL
    assert E0 == pytest.approx(rr.E0, rel=1e-14)


@pytest.mark.parametrize("n,nbins,flags,lag", [
    (30_001, 1024, (True, True, True, True, True), True),    # spectral sums, deposit, resonators
    (4_097, 256, (True, True, True, True, True), False),     # one block (the fused product)
    (10_000, 512, (True, False, False, True, True), True),   # wake without the E0 update
    (5_000, 64, (False, True, True, True, True), False),     # E0 update without the wake
    (2, 2, (True, True, True, True, True), True),            # the smallest grid
    (1_001, 7, (True, True, True, True, True), True),        # an odd bin count (LDS layout)
])
def test_bin_phase_in_one_workgroup_equals_the_launches(dev, n, nbins, flags, lag):
    """A small bunch's bin phase runs in one workgroup (k_bin_whole) unless
    bin_whole = -1 asks for a launch per phase: the same expressions of the
    same operands, so the same bits, monitor included."""
    p = H.SimulationParameters.from_dict(W.ring_3gev_lattice(n_turns=40, flags=flags))
    z0, e0 = O.generate_gaussian(31, 0.0, 4e-3, 0.0, 3e6, n)
    outs = []
    for bw in (0, -1):
        b = H.DeviceBunch(n, dev)
        b.upload(z0, e0)
        kw = dict(resonators=[W.RESONATOR_BB], charge=2e-9) if lag else {}
        r = b.evolve(p, nbins, H.TrackingOptions(seed=3, bin_half_width_sigmas=8.0, monitor=True,
                                                 bin_whole=bw, **kw))
        outs.append((*b.download(), r))
        b.close()
    (za, ea, ra), (zb, eb, rb) = outs
    np.testing.assert_array_equal(za, zb)
    np.testing.assert_array_equal(ea, eb)
    assert (ra.sigma_E, ra.sigma_z, ra.E0, ra.turns_done) == (rb.sigma_E, rb.sigma_z, rb.E0, rb.turns_done)
    np.testing.assert_array_equal(ra.monitor, rb.monitor)
