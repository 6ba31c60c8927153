"""Batched tracking (many bunches per launch) and the serving loop.

A bunch's tiling, reduction order and random stream depend on that bunch
alone, so running it inside a batch or a served mix must give the same bits
as running it alone; these tests hold the batched and served paths to that
(bitwise), plus cancel, the failure hook and HBM-budget admission.
"""
import numpy as np
import pytest

from helpers import COORD_RTOL, W, normwise
from oracle import oracle as O

import haissinskijulia_amd as H
from haissinskijulia_amd._lib import HM_ERR_CANCELLED as CANCELLED

pytestmark = pytest.mark.gpu


def _params(turns, V=3.2e6, flags=(True, True, True, True, True)):
    return H.SimulationParameters.from_dict(W.ring_3gev_lattice(turns, voltage=V, flags=flags))


def _specs():
    """Uneven bunches: sizes, bins, voltages, charges, flags and seeds differ."""
    return [
        dict(n=50_001, nbins=1024, V=3.2e6, charge=1e-9, seed=11, flags=(True, True, True, True, True)),
        dict(n=7_000, nbins=256, V=2.8e6, charge=4e-9, seed=12, flags=(True, True, True, True, True)),
        dict(n=120_000, nbins=4096, V=3.9e6, charge=8e-9, seed=13, flags=(True, True, True, True, True)),
        dict(n=3_333, nbins=512, V=3.0e6, charge=0.0, seed=14, flags=(True, False, True, True, True)),
        dict(n=20_000, nbins=64, V=3.5e6, charge=0.0, seed=15, flags=(False, True, False, True, True)),
    ]


def _opts(s):
    return H.TrackingOptions(seed=s["seed"] + 1000, charge=s["charge"],
                             resonators=[W.RESONATOR_BB] if s["charge"] else (),
                             bin_half_width_sigmas=8.0)


def _bunch(s):
    b = H.DeviceBunch(s["n"])
    b.generate(s["seed"], W.SIGMA_Z, W.SIGMA_E)
    return b


@pytest.mark.parametrize("turns", [1, 40])
def test_batch_equals_each_bunch_alone_bitwise(turns):
    specs = _specs()
    alone = []
    for s in specs:
        b = _bunch(s)
        r = b.evolve(_params(turns, s["V"], s["flags"]), s["nbins"], _opts(s))
        alone.append((r, *b.download()))
        b.close()
    bunches = [_bunch(s) for s in specs]
    rs = H.batch_evolve(bunches, [s["nbins"] for s in specs],
                        [_params(turns, s["V"], s["flags"]) for s in specs],
                        [_opts(s) for s in specs])
    for s, b, r, (ra, za, ea) in zip(specs, bunches, rs, alone):
        z, e = b.download()
        print(f"n={s['n']} nbins={s['nbins']} sigma_E={r.sigma_E:.6e} E0={r.E0:.6e}", flush=True)
        assert np.array_equal(z, za) and np.array_equal(e, ea)
        assert (r.sigma_E, r.sigma_z, r.E0, r.turns_done) == \
               (ra.sigma_E, ra.sigma_z, ra.E0, ra.turns_done)
        b.close()


def test_many_bunch_forms_equal_alone_bitwise():
    # Alone, a bunch of 4096 bins re-reads ~4 MB of spectra per turn and takes
    # the one-kernel inverse phase (k_conv_inv), and its reductions run one
    # workgroup per slice of tile sums.  Sixteen of them in a batch take the
    # spectral sums per frequency chunk (k_conv_spec, k_conv_out) and one
    # workgroup per bunch for the reductions (k_bin_reduce_whole,
    # k_finalize_whole; up to four slices here).  Each form adds the same
    # numbers in the same order: the same bits.
    specs = [dict(n=20_000 + 25_000 * i, nbins=4096, V=2.6e6 + 0.1e6 * i, charge=(1 + i % 8) * 1e-9,
                  seed=40 + i, flags=(True, True, True, True, True)) for i in range(16)]
    alone = []
    for s in specs:
        b = _bunch(s)
        r = b.evolve(_params(30, s["V"]), s["nbins"], _opts(s))
        alone.append((r, *b.download()))
        b.close()
    bunches = [_bunch(s) for s in specs]
    rs = H.batch_evolve(bunches, [s["nbins"] for s in specs], [_params(30, s["V"]) for s in specs],
                        [_opts(s) for s in specs])
    for s, b, r, (ra, za, ea) in zip(specs, bunches, rs, alone):
        z, e = b.download()
        print(f"n={s['n']} sigma_E={r.sigma_E:.6e} E0={r.E0:.6e}", flush=True)
        assert np.array_equal(z, za) and np.array_equal(e, ea)
        assert (r.sigma_E, r.sigma_z, r.E0) == (ra.sigma_E, ra.sigma_z, ra.E0)
        b.close()


def test_batch_of_uneven_turn_counts_matches_alone():
    """Bunches leave the batch at different turns; the others continue."""
    specs = _specs()[:3]
    turns = [5, 17, 11]
    alone = []
    for s, t in zip(specs, turns):
        b = _bunch(s)
        r = b.evolve(_params(t, s["V"]), s["nbins"], _opts(s))
        alone.append((r, *b.download()))
        b.close()
    bunches = [_bunch(s) for s in specs]
    rs = H.batch_evolve(bunches, [s["nbins"] for s in specs],
                        [_params(t, s["V"]) for s, t in zip(specs, turns)],
                        [_opts(s) for s in specs])
    for b, r, t, (ra, za, ea) in zip(bunches, rs, turns, alone):
        z, e = b.download()
        assert r.turns_done == t
        assert np.array_equal(z, za) and np.array_equal(e, ea)
        assert r.E0 == ra.E0
        b.close()


def test_batched_bunch_matches_oracle():
    """One bunch of a batch against the oracle (injected initial state and
    the device's own Philox stream restated by the oracle)."""
    s = _specs()[0]
    turns = 20
    p = W.ring_3gev_lattice(turns, voltage=s["V"])
    z0, e0 = O.generate_gaussian(s["seed"], 0.0, W.SIGMA_Z, 0.0, W.SIGMA_E, s["n"])
    zr, er, rr = O.evolve(z0, e0, s["nbins"], p, seed=s["seed"] + 1000, charge=s["charge"],
                          resonators=[W.RESONATOR_BB], bin_half_width_sigmas=8.0)
    assert rr.status == 0
    target = H.DeviceBunch(s["n"])
    target.upload(z0, e0)
    others = [_bunch(x) for x in _specs()[1:3]]
    bunches = [others[0], target, others[1]]
    specs = [_specs()[1], s, _specs()[2]]
    rs = H.batch_evolve(bunches, [x["nbins"] for x in specs],
                        [_params(turns, x["V"]) for x in specs], [_opts(x) for x in specs])
    z, e = target.download()
    ez, ee = normwise(z, zr), normwise(e, er)
    print(f"normwise z={ez:.2e} dE={ee:.2e}", flush=True)
    assert ez <= COORD_RTOL and ee <= COORD_RTOL
    assert rs[1].E0 == pytest.approx(rr.E0, rel=1e-14)
    for b in bunches:
        b.close()


def _job(s, turns, **kw):
    return H.Job(params=_params(turns, s["V"], s["flags"]), n=s["n"], nbins=s["nbins"],
                 particle_seed=s["seed"], sigma_z=W.SIGMA_Z, sigma_E=W.SIGMA_E, options=_opts(s), **kw)


def test_served_jobs_equal_alone_runs_and_finish_in_order_of_work():
    specs = _specs()
    turns = [30, 12, 25, 7, 18]
    srv = H.Server(budget_bytes=8 << 30, max_slots=8)
    ids = [srv.submit(_job(s, t)) for s, t in zip(specs, turns)]
    res = {r.job_id: r for r in srv.run(turns_per_tick=5)}
    assert sorted(res) == sorted(ids)
    for s, t, jid in zip(specs, turns, ids):
        b = _bunch(s)
        ra = b.evolve(_params(t, s["V"], s["flags"]), s["nbins"], _opts(s))
        z, e = b.download()
        b.close()
        r = res[jid]
        assert r.status == 0
        assert r.result.turns_done == t
        assert (r.result.sigma_E, r.result.sigma_z, r.result.E0) == (ra.sigma_E, ra.sigma_z, ra.E0)
        assert r.centroid_z == pytest.approx(float(np.mean(z)), rel=1e-9, abs=1e-15)
        assert r.t_submit <= r.t_admit <= r.t_done
    srv.close()


def test_cancel_and_failure_hook():
    specs = _specs()
    srv = H.Server(budget_bytes=8 << 30, max_slots=2)
    a = srv.submit(_job(specs[0], 40))
    b = srv.submit(_job(specs[1], 40, fail_after_turns=10))
    c = srv.submit(_job(specs[2], 40))     # queued behind the two slots
    d = srv.submit(_job(specs[3], 15))
    srv.tick(5)
    srv.cancel(a)                          # resident
    srv.cancel(c)                          # may be queued or resident by now
    res = {r.job_id: r for r in srv.run(turns_per_tick=5)}
    assert res[a].status == CANCELLED and res[a].result.turns_done < 40
    assert res[b].status == CANCELLED and res[b].result.turns_done == 10
    assert res[c].status == CANCELLED
    assert res[d].status == 0 and res[d].result.turns_done == 15
    srv.close()


def test_budget_admits_one_job_at_a_time():
    specs = _specs()[:3]
    # the jobs need about 1.2, 0.5 and 2.6 MB (particles, bin tables, FFT
    # work); 1.5 MB holds no two of them, and the 2.6 MB one runs alone
    srv = H.Server(budget_bytes=1_500_000, max_slots=8)
    ids = [srv.submit(_job(s, 6)) for s in specs]
    res = [r for r in srv.run(turns_per_tick=3)]
    order = {r.job_id: r for r in res}
    adm = sorted((order[i].t_admit, order[i].t_done) for i in ids)
    for (a0, d0), (a1, _) in zip(adm, adm[1:]):
        assert a1 >= d0 - 1e-9          # the next job starts after the previous one ends
    assert all(order[i].status == 0 for i in ids)
    srv.close()


def test_pair_stats_match_numpy():
    rng = np.random.default_rng(3)
    z = rng.normal(1e-4, 4e-3, 300_001)
    e = 2e5 * (z / 4e-3) + rng.normal(-50.0, 3e6, z.size)
    s = H.pair_stats(z, e)
    assert s["mean_z"] == pytest.approx(z.mean(), rel=1e-12)
    assert s["std_z"] == pytest.approx(z.std(ddof=1), rel=1e-12)
    assert s["std_dE"] == pytest.approx(e.std(ddof=1), rel=1e-12)
    assert s["cor"] == pytest.approx(np.corrcoef(z, e)[0, 1], rel=1e-11)


def test_emittance_figure_of_merit():
    """compute_fom (figures_of_merit.jl:36-47) from the resident run's own
    correlation and from the downloaded particles agree with numpy."""
    import dataclasses
    s = _specs()[0]
    b = _bunch(s)
    r0 = b.evolve(_params(30, s["V"]), s["nbins"], _opts(s))
    assert np.isnan(r0.corr_z_dE)            # not asked for: no extra reads of the bunch
    with pytest.raises(ValueError):
        H.compute_fom(H.EmittanceFoM(), None, r0)
    b.close()
    b = _bunch(s)
    r = b.evolve(_params(30, s["V"]), s["nbins"], dataclasses.replace(_opts(s), fom_stats=True))
    assert (r.sigma_E, r.sigma_z, r.E0) == (r0.sigma_E, r0.sigma_z, r0.E0)
    z, e = b.download()
    b.close()
    assert r.mean_z == pytest.approx(z.mean(), rel=1e-12)
    ref = r.sigma_z * r.sigma_E * np.sqrt(1 - np.corrcoef(z, e)[0, 1] ** 2)
    f1 = H.compute_fom(H.EmittanceFoM(), None, r)
    f2 = H.compute_fom(H.EmittanceFoM(), H.Particles(z, e), (r.sigma_E, r.sigma_z, r.E0))
    assert f1 == pytest.approx(ref, rel=1e-10) and f2 == pytest.approx(ref, rel=1e-10)
    assert H.compute_fom(H.EnergySpreadFoM(), None, r) == r.sigma_E
    assert H.compute_fom(H.BunchLengthFoM(), None, r) == r.sigma_z


def test_served_voltage_pair_is_a_common_random_number_difference():
    """A sensitivity request runs V (1 -+ 0.5%) with the same seeds; served, its
    members equal the two runs done alone, and the FD derivative follows."""
    s = _specs()[2]
    srv = H.Server(budget_bytes=8 << 30, max_slots=8)
    other = srv.submit(_job(_specs()[1], 25))
    lo, hi = srv.submit_voltage_pair(_job(s, 25), 0.005)
    res = {r.job_id: r for r in srv.run(turns_per_tick=5)}
    srv.close()
    assert res[other].status == 0
    import dataclasses
    alone = []
    for f in (1 - 0.005, 1 + 0.005):
        b = _bunch(s)
        p = _params(25, s["V"]).replace(voltage=s["V"] * f)
        alone.append(b.evolve(p, s["nbins"], dataclasses.replace(_opts(s), fom_stats=True)))
        b.close()
    for jid, ra in zip((lo, hi), alone):
        assert (res[jid].result.sigma_E, res[jid].result.E0) == (ra.sigma_E, ra.E0)
    for fom in (H.EnergySpreadFoM(), H.BunchLengthFoM(), H.EmittanceFoM()):
        d = H.fd_sensitivity(fom, res[lo].result, res[hi].result, s["V"], 0.005)
        da = H.fd_sensitivity(fom, alone[0], alone[1], s["V"], 0.005)
        assert np.isfinite(d) and d == da


def test_failed_admission_releases_its_slots():
    """Two jobs admitted in one tick, the second failing its set-up inside the
    batch after its slot is taken (the fail_setup hook): the batch gives every
    slot of that admission back, the healthy job is admitted again alone and
    ends with the bits it has alone; a later job runs normally."""
    import dataclasses
    s0, s1 = _specs()[0], _specs()[1]
    srv = H.Server(budget_bytes=8 << 30, max_slots=8)
    good = srv.submit(_job(s0, 20))
    bad = srv.submit(dataclasses.replace(_job(s1, 20), fail_setup=True))
    res = {r.job_id: r for r in srv.run(turns_per_tick=5)}
    later = srv.submit(_job(s1, 20))
    res.update({r.job_id: r for r in srv.run(turns_per_tick=5)})
    srv.close()
    assert res[bad].status == CANCELLED
    assert res[good].status == 0 and res[later].status == 0
    for jid, s in ((good, s0), (later, s1)):
        b = _bunch(s)
        ra = b.evolve(_params(20, s["V"], s["flags"]), s["nbins"], _opts(s))
        b.close()
        assert (res[jid].result.sigma_E, res[jid].result.sigma_z, res[jid].result.E0) == \
               (ra.sigma_E, ra.sigma_z, ra.E0)


def test_tile_knobs_keep_bits_alone_and_batched():
    """tile_pairs sets a bunch's tiling, hence the order of its per-tile sums:
    under the same value a bunch gives the same bits alone and batched, and
    every value stays within the parity bar of the rule's result (the oracle's
    distance to either is rounding).  run_pairs only moves the pass's
    workgroup boundaries: the same bits whatever its value."""
    specs = [dict(n=300_001, nbins=1024, V=3.2e6, charge=3e-9, seed=71, flags=(True,) * 5),
             dict(n=40_000, nbins=4096, V=3.6e6, charge=6e-9, seed=72, flags=(True,) * 5)]
    turns = 20
    runs = {}
    for tp, rp in ((0, 0), (3, 0), (3, 1), (3, 1 << 20), (16, 0)):
        bunches = [_bunch(s) for s in specs]
        opts = []
        for s in specs:
            o = _opts(s)
            o.tile_pairs, o.run_pairs = tp, rp
            opts.append(o)
        rs = H.batch_evolve(bunches, [s["nbins"] for s in specs],
                            [_params(turns, s["V"], s["flags"]) for s in specs], opts)
        runs[(tp, rp)] = [(r.E0, *b.download()) for r, b in zip(rs, bunches)]
        for b in bunches:
            b.close()
        print(f"tile_pairs {tp} run_pairs {rp}: E0 {[r.E0 for r in rs]}", flush=True)
    # alone under tile_pairs = 3
    s = specs[0]
    b = _bunch(s)
    o = _opts(s)
    o.tile_pairs = 3
    r = b.evolve(_params(turns, s["V"], s["flags"]), s["nbins"], o)
    za, ea = b.download()
    b.close()
    E0b, zb, eb = runs[(3, 0)][0]
    assert r.E0 == E0b and np.array_equal(za, zb) and np.array_equal(ea, eb)
    for rp in (1, 1 << 20):   # the run cost changes nothing
        for (E0x, zx, ex), (E0y, zy, ey) in zip(runs[(3, 0)], runs[(3, rp)]):
            assert E0x == E0y and np.array_equal(zx, zy) and np.array_equal(ex, ey)
    for tp in (3, 16):        # another tiling: another order of the sums, the same physics
        for (E0x, zx, ex), (E0y, zy, ey) in zip(runs[(0, 0)], runs[(tp, 0)]):
            assert E0y == pytest.approx(E0x, rel=1e-14)
            assert normwise(zy, zx) <= COORD_RTOL and normwise(ey, ex) <= COORD_RTOL


def test_the_tuners_rule_is_the_librarys_rule(monkeypatch):
    """tune.rule_tile_pairs restates the library's default tiling: with no
    tuned entry, tile_pairs = 0 and tile_pairs = rule_tile_pairs(n) give the
    same bits in the state, the per-turn monitor and the final moments.  A
    neighbouring tiling reorders the floating-point sums of the per-tile
    moments: the moments change in their last bits, while the state of this
    short run does not (the histogram and the cloud-in-cell E0 sums are
    integer)."""
    from haissinskijulia_amd import tune
    monkeypatch.setattr(tune, "_table", {"fft_block": {}, "tile_pairs": {}, "run_pairs": {}})
    s = dict(n=3_000_001, nbins=2048, V=3.2e6, charge=5e-9, seed=73, flags=(True,) * 5)
    rule = tune.rule_tile_pairs(s["n"])
    assert rule == 2        # ceil(ceil(1500001 / 1024) / 1024)
    out = {}
    for tp in (0, rule, rule + 1):
        b = _bunch(s)
        o = _opts(s)
        o.tile_pairs, o.monitor = tp, True
        r = b.evolve(_params(10, s["V"], s["flags"]), s["nbins"], o)
        out[tp] = [np.array([r.E0, r.sigma_E, r.sigma_z, r.mean_z, r.mean_dE]),
                   np.asarray(r.monitor), *b.download()]
        b.close()
    same = lambda x, y: [bool(np.array_equal(u, v)) for u, v in zip(x, y)]
    assert same(out[0], out[rule]) == [True] * 4
    other = same(out[0], out[rule + 1])
    print(f"tile_pairs {rule + 1} against the rule: scalars, monitor, z, dE equal: {other}",
          flush=True)
    # the comparison can tell tilings apart: here the moments (sigma_z, sigma_E
    # and the monitor's sigma_z) differ in their last bits
    assert not (other[0] and other[1])


def test_large_tiles_of_a_16384_bin_bunch(monkeypatch):
    """64 pairs per thread per tile, the tuned tiling of the 1e8 workload, on
    a 1.2e7-particle bunch with 16384 bins (the E0 update's interpolation
    sums, as at 1e8; 92 tiles of 65536 pairs against the rule's 977 of
    6144): within the parity bar of the rule's result, and the same bits
    alone and batched with a second bunch."""
    from haissinskijulia_amd import tune
    monkeypatch.setattr(tune, "_table", {"fft_block": {}, "tile_pairs": {}, "run_pairs": {}})
    s = dict(n=12_000_001, nbins=16384, V=3.4e6, charge=10e-9, seed=74, flags=(True,) * 5)
    small = dict(n=60_000, nbins=1024, V=3.0e6, charge=2e-9, seed=75, flags=(True,) * 5)
    turns = 12
    out = {}
    for tp in (0, 64):
        b = _bunch(s)
        o = _opts(s)
        o.tile_pairs = tp
        r = b.evolve(_params(turns, s["V"], s["flags"]), s["nbins"], o)
        out[tp] = (r, *b.download())
        b.close()
        print(f"tile_pairs {tp}: E0 {r.E0!r} sigma_z {r.sigma_z!r}", flush=True)
    (r0, z0, e0), (r64, z64, e64) = out[0], out[64]
    assert r64.E0 == pytest.approx(r0.E0, rel=1e-14)
    assert normwise(z64, z0) <= COORD_RTOL and normwise(e64, e0) <= COORD_RTOL
    bunches = [_bunch(small), _bunch(s)]
    opts = [_opts(small), _opts(s)]
    opts[1].tile_pairs = 64
    rs = H.batch_evolve(bunches, [small["nbins"], s["nbins"]],
                        [_params(turns, x["V"], x["flags"]) for x in (small, s)], opts)
    zb, eb = bunches[1].download()
    for b in bunches:
        b.close()
    assert rs[1].E0 == r64.E0 and np.array_equal(zb, z64) and np.array_equal(eb, e64)


def test_closing_a_device_releases_its_bunches_and_servers_first():
    dev = H.MI355X(0)
    b = H.DeviceBunch(1000, dev)
    b.generate(3, W.SIGMA_Z, W.SIGMA_E)
    srv = H.Server(1 << 30, 2, dev)
    dev.close()
    assert b._h is None and srv._h is None and dev._ctx is None
    b.close()
    srv.close()
    dev.close()
