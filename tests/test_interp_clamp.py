"""The turn pass clamps the interpolation coordinate u = (z - c0) * inv to
[0, u1], u1 = (c1 - c0) * inv, where the reference clamps z to [c0, c1]
(src/wakefield.jl:64-78) and then forms u (CPU only).

Both are IEEE FP64 expressions evaluated with separate roundings (the
library is built with -ffp-contract=off), so numpy evaluates them as the
device does.  The claim checked here: for every z that is not NaN the two
give the same u bit for bit, hence the same cell and weight, on the
reference's own bin grids and on z at and around the ends.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _grid(zw, nb):
    _, centers, _ = O.bin_grid(zw, nb)
    c0, c1 = centers[0], centers[-1]
    inv = 1.0 / ((c1 - c0) / float(nb - 1))   # wakefield.jl:39-42
    return c0, c1, inv


def _u_reference(z, c0, c1, inv):
    cl = np.minimum(np.maximum(z, c0), c1)
    return (cl - c0) * inv


def _u_pass(z, c0, c1, inv):
    # fmax/fmin: a NaN operand gives the other one, as the device's v_max_f64
    # and v_min_f64 do
    u1 = (c1 - c0) * inv
    return np.fmin(np.fmax((z - c0) * inv, 0.0), u1)


def _probe_points(c0, c1, rng):
    step = (c1 - c0)
    z = [c0, c1, -np.inf, np.inf, 0.0, -0.0, 1e300, -1e300, 5e-324, -5e-324]
    for e in (c0, c1):
        x = e
        for _ in range(40):   # the doubles just below and above each end
            x = np.nextafter(x, -np.inf)
            z.append(x)
        x = e
        for _ in range(40):
            x = np.nextafter(x, np.inf)
            z.append(x)
    z.extend(rng.uniform(c0 - 0.5 * step, c1 + 0.5 * step, 200_000))
    z.extend(rng.normal(0.0, step / 6, 200_000))
    return np.array(z, dtype=np.float64)


@pytest.mark.parametrize("zw,nb", [(0.03, 2), (0.03, 3), (0.032, 1024), (0.0123456, 4096),
                                   (0.064, 5087), (0.25, 16384), (1e-6, 8192), (7.5e-3, 65536)])
def test_clamped_coordinate_equals_the_reference_clamp(zw, nb):
    c0, c1, inv = _grid(zw, nb)
    z = _probe_points(c0, c1, np.random.default_rng(nb))
    ur = _u_reference(z, c0, c1, inv)
    up = _u_pass(z, c0, c1, inv)
    assert np.array_equal(ur.view(np.int64), up.view(np.int64))   # bit for bit, +0 included
    k = np.minimum(np.floor(up).astype(np.int64), nb - 2)
    assert k.min() >= 0 and k.max() <= nb - 2


def test_a_nan_z_lands_on_cell_zero():
    c0, c1, inv = _grid(0.032, 1024)
    u = _u_pass(np.array([np.nan]), c0, c1, inv)
    assert u[0] == 0.0 and not np.signbit(u[0])


def test_a_planted_clamp_of_z_to_the_wrong_end_is_caught():
    # the check itself: an upper bound one rounding below u1 must differ from
    # the reference for z at and beyond c1
    c0, c1, inv = _grid(0.032, 1024)
    z = np.array([c1, np.nextafter(c1, np.inf), 1.0])
    wrong = np.fmin(np.fmax((z - c0) * inv, 0.0), (c1 - c0) * inv * (1 - 2 ** -52))
    assert not np.array_equal(_u_reference(z, c0, c1, inv), wrong)
