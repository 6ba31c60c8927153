"""C ABI of libhaissinski_mi355x.so without a GPU: the library loads, exports
every function include/haissinski_mi355x.h declares, and its host-only bin
grid helper is correctly rounded.  No kernel is launched here."""
import ctypes as C
import re
from fractions import Fraction

import numpy as np
import pytest

from haissinskijulia_amd import _lib
from haissinskijulia_amd import api


def test_library_loads_and_exports_every_declared_symbol():
    L = _lib.load()
    names = _lib.header_symbols()
    assert len(names) >= 29
    for name in names:
        assert hasattr(L, name), name
    # and the header declares what the binding binds
    text = open(_lib.HEADER).read()
    assert "extern \"C\"" in text
    assert all(re.search(rf"\b{n}\s*\(", text) for n in names)


def test_struct_layouts_match_the_c_compiler(tmp_path):
    # the header compiled by gcc (tests/abi_layout.c) against the ctypes binding:
    # every struct's size and every field's offset
    import os
    import subprocess
    here = os.path.dirname(os.path.abspath(__file__))
    exe = tmp_path / "abi_layout"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-o", str(exe), os.path.join(here, "abi_layout.c")],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    binding = {"hm_params": _lib.HmParams, "hm_options": _lib.HmOptions, "hm_result": _lib.HmResult,
               "hm_job": _lib.HmJob, "hm_job_result": _lib.HmJobResult}
    seen = 0
    for line in out.splitlines():
        st, field, val = line.split()
        cls = binding[st]
        if field == "sizeof":
            assert C.sizeof(cls) == int(val), st
        else:
            assert getattr(cls, field).offset == int(val), (st, field)
        seen += 1
    assert seen >= 60


@pytest.mark.parametrize("zw,nb", [(0.03, 1024), (7.5 * 4.123456789e-3, 1000), (1e-6, 7),
                                   (0.0123, 1), (0.05, 16384), (3.3e-2, 8191)])
def test_host_bin_grid_is_correctly_rounded(zw, nb):
    edges, centers, step = api.bin_grid(zw, nb)
    fz = Fraction(zw)
    for k in list(range(0, nb + 1, max(1, nb // 101))) + [nb]:
        assert edges[k] == float(fz * (2 * k - nb) / nb), k
    for i in list(range(0, nb, max(1, nb // 103))) + [nb - 1]:
        assert centers[i] == float(fz * (2 * i + 1 - nb) / nb), i
    assert step == float(2 * fz / nb)


def test_host_bin_grid_matches_oracle():
    from oracle import oracle as O
    for zw, nb in [(0.0301, 999), (0.017, 4096)]:
        e1, c1, s1 = api.bin_grid(zw, nb)
        e2, c2, s2 = O.bin_grid(zw, nb)
        np.testing.assert_array_equal(e1, e2)
        np.testing.assert_array_equal(c1, c2)
        assert s1 == s2


def test_bad_arguments_are_rejected_on_the_host():
    with pytest.raises(_lib.HaissinskiError) as ei:
        api.bin_grid(-1.0, 10)
    assert ei.value.status == _lib.HM_ERR_ARG
    with pytest.raises(_lib.HaissinskiError) as ei:
        api.create_simulation_buffers(10, 0)
    assert ei.value.status == _lib.HM_ERR_NBINS


def test_host_helpers_follow_reference():
    assert api.next_power_of_two(0) == 1 and api.next_power_of_two(1) == 1
    assert api.next_power_of_two(1000) == 1024 and api.next_power_of_two(2048) == 2048
    assert api.z_to_ϕ(0.001, 10.0, 2.5) == -(0.001 * 10.0 - 2.5)
    assert api.ϕ_to_z(api.z_to_ϕ(0.001, 10.0, 2.5), 10.0, 2.5) == pytest.approx(0.001, rel=1e-12)
    assert api.calc_rf_factor(5e8, 0.0) == float("inf")
    p = api.SimulationParameters.from_dict({"E0": 1.0, "mass": 1.0, "voltage": 1.0, "harmonic": 1,
                                            "radius": 1.0, "pipe_radius": 1.0, "α_c": 2.0, "ϕs": 3.0,
                                            "freq_rf": 1.0, "n_turns": 1, "use_wakefield": True,
                                            "update_η": False, "update_E0": True, "SR_damping": True,
                                            "use_excitation": True})
    assert p.alpha_c == 2.0 and p.phi_s == 3.0 and p.update_eta is False


def test_context_creation_fails_loudly_without_gfx950():
    L = _lib.load()
    h = C.c_void_p()
    st = L.hm_create(0, C.byref(h))
    if st == 0:
        L.hm_destroy(h)
        pytest.skip("a gfx950 device is present")
    assert st in (_lib.HM_ERR_DEVICE, _lib.HM_ERR_ARG)
    assert L.hm_last_error()
