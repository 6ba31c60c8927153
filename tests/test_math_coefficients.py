"""The polynomial constants of the kernels' elementary functions
(haissinskijulia_amd/csrc/hm_math.hpp) re-derived from their definition:
interpolants at the Chebyshev nodes (first kind) of the remainder after the
leading Taylor term, in 60-digit arithmetic (mpmath), on the reduced
interval of each function:

* sin r = r - r^3 p(r^2), p(t) = (r - sin r)/r^3, t in [0, (pi/4)^2], degree 6;
* cos r = 1 - r^2/2 + r^4 p(r^2), p(t) = (cos r - 1 + t/2)/t^2, same interval,
  degree 5;
* log m = 2 atanh s = 2s (1 + s^2 p(s^2)), p(t) = (atanh(s)/s - 1)/t,
  s = (m-1)/(m+1) for m in [sqrt(1/2), sqrt(2)), t in [0, s_max^2], degree 7.

Every constant of the header must be the double nearest to the re-derived
coefficient.  The RF kick's sin (src/rf_kick.jl:85-86) and the Box-Muller
transform of the excitation draws use them.
"""
import os
import re

import mpmath as mp
import pytest

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "haissinskijulia_amd", "csrc", "hm_math.hpp")


def _chebyshev_interpolant(f, a, b, degree):
    n = degree
    xs = [(a + b) / 2 + (b - a) / 2 * mp.cos((2 * k + 1) * mp.pi / (2 * (n + 1)))
          for k in range(n + 1)]
    A = mp.matrix([[x ** j for j in range(n + 1)] for x in xs])
    c = mp.lu_solve(A, mp.matrix([f(x) for x in xs]))
    return [c[j] for j in range(n + 1)]


def _fsin(t):
    r = mp.sqrt(t)
    return (r - mp.sin(r)) / r ** 3


def _fcos(t):
    return (mp.cos(mp.sqrt(t)) - 1 + t / 2) / t ** 2


def _flog(t):
    s = mp.sqrt(t)
    return (mp.atanh(s) / s - 1) / t


def _header_polys():
    """{function: [constant term, ..., highest]} from the Horner chains."""
    text = open(HEADER).read()
    out = {}
    for name in ("sin_poly", "cos_poly", "log_pos"):
        body = text[text.index(name):]
        body = body[:body.index("\n}\n")]
        first = re.search(r"double p = ([-+0-9.e]+);", body).group(1)
        rest = re.findall(r"p = HM_FMAK\(p, \w+, ([-+0-9.e]+)\);", body)
        out[name] = [float(x) for x in reversed([first] + rest)]
    return out


@pytest.mark.parametrize("name,f,upper,degree", [
    ("sin_poly", _fsin, lambda: (mp.pi / 4) ** 2, 6),
    ("cos_poly", _fcos, lambda: (mp.pi / 4) ** 2, 5),
    ("log_pos", _flog, lambda: ((mp.sqrt(2) - 1) / (mp.sqrt(2) + 1)) ** 2, 7),
])
def test_header_constants_are_the_chebyshev_interpolants(name, f, upper, degree):
    mp.mp.dps = 60
    coef = _chebyshev_interpolant(f, mp.mpf(0), upper(), degree)
    got = _header_polys()[name]
    assert len(got) == degree + 1
    for j, (c, g) in enumerate(zip(coef, got)):
        assert float(c) == g, (name, j, mp.nstr(c, 25), g)


def test_interpolants_meet_the_stated_truncation_error():
    # the interpolation error on the reduced interval, relative to the value
    # the polynomial feeds (r^3 p for sin, r^4 p for cos, s^3 p for log), is far
    # below an ulp of the function
    mp.mp.dps = 60
    for f, up, deg, scale in ((_fsin, (mp.pi / 4) ** 2, 6, lambda t: mp.sqrt(t) ** 3),
                              (_fcos, (mp.pi / 4) ** 2, 5, lambda t: t ** 2),
                              (_flog, ((mp.sqrt(2) - 1) / (mp.sqrt(2) + 1)) ** 2, 7,
                               lambda t: 2 * mp.sqrt(t) ** 3)):
        c = _chebyshev_interpolant(f, mp.mpf(0), up, deg)
        worst = mp.mpf(0)
        for k in range(1, 400):
            t = up * k / 400
            p = sum(cj * t ** j for j, cj in enumerate(c))
            worst = max(worst, abs(p - f(t)) * scale(t))
        assert worst < mp.mpf("1e-18")
