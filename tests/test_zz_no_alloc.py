"""hm_evolve_device_noalloc (the allocation-free resident evolve) against
hm_bunch_evolve, through the C ABI: two calls in a row on the same GPU, the
second with a smaller bunch and bin count, each compared with the allocating
entry point on an identical bunch (the same kernels: the same bits)."""
import ctypes as C

import numpy as np
import pytest

from helpers import W

import haissinskijulia_amd as H
from haissinskijulia_amd import _lib
from haissinskijulia_amd._lib import HM_ERR_ARG, HM_ERR_LIMIT, HmOptions, HmResult

pytestmark = pytest.mark.gpu


def _noalloc(bunch, nbins, params, opts):
    L = _lib.load()
    zp, ep = C.c_void_p(), C.c_void_p()
    _lib.check(L.hm_bunch_device_ptrs(bunch._h, C.byref(zp), C.byref(ep)))
    keep = []
    o = opts._c(keep, bunch.n, params.n_turns, nbins)
    r = HmResult()
    st = L.hm_evolve_device_noalloc(bunch.device.ctx, zp, ep, bunch.n, nbins, C.byref(params._c()),
                                    C.byref(o), C.byref(r))
    return st, r


CASES = [
    # the headline's bins, wake and flags on a 2e6-particle bunch
    dict(n=2_000_001, nbins=16384, turns=30, seed=11, V=3.2e6,
         flags=(True, True, True, True, True), charge=10e-9,
         resonators=[W.RESONATOR_BB, W.RESONATOR_LOWQ]),
    # then a smaller one: fewer particles and bins, constant eta, FoM statistics
    dict(n=150_000, nbins=1024, turns=45, seed=12, V=2.9e6,
         flags=(True, False, True, True, True), charge=2e-9, resonators=[W.RESONATOR_BB],
         fom_stats=True),
]


def _opts(c):
    return H.TrackingOptions(seed=c["seed"] + 5, resonators=c["resonators"], charge=c["charge"],
                             bin_half_width_sigmas=8.0, fom_stats=c.get("fom_stats", False))


def test_noalloc_equals_allocating_evolve_on_two_calls():
    dev = H.MI355X.default()
    for c in CASES:
        print(f"noalloc case n={c['n']} nbins={c['nbins']} turns={c['turns']}", flush=True)
        p = H.SimulationParameters.from_dict(W.ring_3gev_lattice(c["turns"], voltage=c["V"],
                                                                 flags=c["flags"]))
        a = H.DeviceBunch(c["n"], dev)
        b = H.DeviceBunch(c["n"], dev)
        a.generate(c["seed"], W.SIGMA_Z, W.SIGMA_E)
        b.generate(c["seed"], W.SIGMA_Z, W.SIGMA_E)
        st, r = _noalloc(a, c["nbins"], p, _opts(c))
        assert st == 0, _lib.load().hm_last_error()
        rb = b.evolve(p, c["nbins"], _opts(c))
        za, ea = a.download()
        zb, eb = b.download()
        a.close()
        b.close()
        np.testing.assert_array_equal(za, zb)
        np.testing.assert_array_equal(ea, eb)
        got = (r.sigma_E, r.sigma_z, r.E0, r.turns_done, r.sigma_E0, r.sigma_z0, r.mean_z, r.mean_dE)
        want = (rb.sigma_E, rb.sigma_z, rb.E0, rb.turns_done, rb.sigma_E0, rb.sigma_z0, rb.mean_z,
                rb.mean_dE)
        assert got == want
        if c.get("fom_stats"):
            assert r.corr_z_dE == rb.corr_z_dE and np.isfinite(r.corr_z_dE)
        else:
            assert np.isnan(r.corr_z_dE) and np.isnan(rb.corr_z_dE)


def test_noalloc_rejects_what_it_cannot_hold():
    dev = H.MI355X.default()
    a = H.DeviceBunch(1000, dev)
    a.generate(1, W.SIGMA_Z, W.SIGMA_E)
    p = H.SimulationParameters.from_dict(W.ring_3gev_lattice(3))
    st, _ = _noalloc(a, 65536, p, H.TrackingOptions(resonators=[W.RESONATOR_BB], charge=1e-9))
    assert st == HM_ERR_LIMIT
    st, _ = _noalloc(a, 256, p, H.TrackingOptions(resonators=[W.RESONATOR_BB], charge=1e-9,
                                                   normals=np.zeros((3, 1000))))
    assert st == HM_ERR_ARG
    st, _ = _noalloc(a, 256, p, H.TrackingOptions(resonators=[W.RESONATOR_BB], charge=1e-9,
                                                   monitor=True))
    assert st == HM_ERR_ARG
    a.close()
