"""The kernels' elementary functions (haissinskijulia_amd/csrc/hm_math.hpp),
compiled for the CPU from the same source, against numpy's libm (CPU)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("math") / "libhmmath.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                    os.path.join(HERE, "math_host_shim.cpp"), "-o", str(out)], check=True)
    L = C.CDLL(str(out))
    dp = C.POINTER(C.c_double)
    for name, nargs in (("t_sin", 2), ("t_log", 2), ("t_rcp", 2), ("t_sincos", 3)):
        getattr(L, name).argtypes = [dp] * nargs + [C.c_int64]
    return L


def _call(L, name, *arrays):
    ps = [a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrays]
    getattr(L, name)(*ps, arrays[0].size)


def _ulp_err(got, ref):
    return np.abs(got - ref) / np.spacing(np.abs(ref))


def test_sin_cos_within_two_ulp(lib):
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-8, 8, 200_000), rng.uniform(2.0, 3.6, 200_000),
                        np.linspace(-1e4, 1e4, 10_001), [0.0, np.pi, np.pi / 2, -np.pi / 4]])
    y = np.empty_like(x)
    _call(lib, "t_sin", x, y)
    s, c = np.empty_like(x), np.empty_like(x)
    _call(lib, "t_sincos", x, s, c)
    np.testing.assert_array_equal(y, s)
    ref_s, ref_c = np.sin(x), np.cos(x)
    # ulp error away from zeros of the function, absolute error near them
    big_s, big_c = np.abs(ref_s) > 1e-3, np.abs(ref_c) > 1e-3
    assert np.max(_ulp_err(s[big_s], ref_s[big_s])) <= 2.0
    assert np.max(_ulp_err(c[big_c], ref_c[big_c])) <= 2.0
    assert np.max(np.abs(s - ref_s)) <= 4e-16 * max(1.0, np.max(np.abs(x)) * 1e-4)
    assert np.max(np.abs(c - ref_c)) <= 4e-16 * max(1.0, np.max(np.abs(x)) * 1e-4)


def test_box_muller_angles(lib):
    u = np.arange(0, 2 ** 20) * 2.0 ** -20
    th = (2 * np.pi) * u
    s, c = np.empty_like(th), np.empty_like(th)
    _call(lib, "t_sincos", th, s, c)
    assert np.max(np.abs(s - np.sin(th))) <= 3e-16
    assert np.max(np.abs(c - np.cos(th))) <= 3e-16


def test_log_within_two_ulp(lib):
    rng = np.random.default_rng(1)
    u = np.concatenate([rng.uniform(0, 1, 200_000), 2.0 ** -rng.uniform(0, 53, 100_000),
                        np.array([2.0 ** -53, 0.5, 1.0, 1 - 2 ** -53, np.sqrt(0.5), 0.70710678118654757]),
                        rng.uniform(1, 1e6, 10_000)])
    u = u[u > 0]
    y = np.empty_like(u)
    _call(lib, "t_log", u, y)
    ref = np.log(u)
    nz = np.abs(ref) > 1e-300
    assert np.max(_ulp_err(y[nz], ref[nz])) <= 2.0
    assert y[u == 1.0][0] == 0.0


def test_rcp(lib):
    x = np.random.default_rng(2).uniform(1.0, 1e8, 100_000)
    y = np.empty_like(x)
    _call(lib, "t_rcp", x, y)
    assert np.max(_ulp_err(y, 1.0 / x)) <= 1.0
