"""Pins of the oracle's cubic spline (SmoothedParticles.h:60-72,
SmoothedParticles.cpp:13-23, 74-80) against the mathematics: closed-form
normalisation constants, unit integral, derivative = finite difference of
the value, support and continuity conditions, and the packed-lattice sum."""
import math

import numpy as np
import pytest
from scipy import integrate

import oracle_py as O


@pytest.mark.parametrize("dim", [2, 3])
def test_normalisation_constant(dim):
    r = 0.0125
    h = 4 * r
    c0 = 40.0 / (7 * math.pi * h * h) if dim == 2 else 8.0 / (math.pi * h ** 3)
    assert O.kernel(dim, r, 0.0) == pytest.approx(c0, rel=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_unit_integral(dim):
    r = 0.01
    h = 4 * r
    shell = (lambda s: 2 * math.pi * s) if dim == 2 else (lambda s: 4 * math.pi * s * s)
    val, _ = integrate.quad(lambda s: O.kernel(dim, r, s) * shell(s), 0, h,
                            points=[h / 2], epsabs=1e-14, epsrel=1e-13)
    assert val == pytest.approx(1.0, abs=1e-10)


@pytest.mark.parametrize("dim", [2, 3])
def test_derivative_is_slope(dim):
    r = 0.02
    h = 4 * r
    for x in np.linspace(0.05, 0.95, 19):
        if abs(x - 0.5) < 0.02:
            continue
        s = x * h
        e = 1e-6 * h
        fd = (O.kernel(dim, r, s + e) - O.kernel(dim, r, s - e)) / (2 * e)
        assert O.kernel_d1(dim, r, s) == pytest.approx(fd, rel=1e-6, abs=1e-9 * O.kernel(dim, r, 0) / h)


@pytest.mark.parametrize("dim", [2, 3])
def test_support_and_continuity(dim):
    r = 0.01
    h = 4 * r
    assert O.kernel(dim, r, h) == 0.0
    assert O.kernel_d1(dim, r, h) == 0.0
    assert O.kernel(dim, r, 1.5 * h) == 0.0
    assert O.kernel_d1(dim, r, 0.0) == 0.0
    e = 1e-9 * h
    assert O.kernel(dim, r, h / 2 - e) == pytest.approx(O.kernel(dim, r, h / 2 + e), rel=1e-7)
    assert O.kernel_d1(dim, r, h / 2 - e) == pytest.approx(O.kernel_d1(dim, r, h / 2 + e), rel=1e-7)
    # value at x = 1/2 from the closed form 2 c0 (1/2)^3 = c0 / 4
    assert O.kernel(dim, r, h / 2) == pytest.approx(O.kernel(dim, r, 0) / 4, rel=1e-14)
    # minimum of W' at x = 1/3: c1 * (1/3) * (1 - 2) = -c1 / 3, c1 = 6 c0 / h
    assert O.kernel_d1(dim, r, h / 3) == pytest.approx(-2 * O.kernel(dim, r, 0) / h, rel=1e-13)


def _w(dim, r, d):
    h = 4 * r
    c0 = 40.0 / (7 * math.pi * h * h) if dim == 2 else 8.0 / (math.pi * h ** 3)
    x = d / h
    if x < 0.5:
        return c0 * (6 * x * x * (x - 1) + 1)
    if x < 1:
        return 2 * c0 * (1 - x) ** 3
    return 0.0


@pytest.mark.parametrize("dim", [2, 3])
def test_packed_kernel_sum(dim):
    r = 0.01
    # 2D: hexagonal lattice of spacing 2r, 6 + 6 neighbours at 2r, 2 sqrt3 r;
    # 3D: 12 at 2r, 6 at 2 sqrt2 r, 24 at 2 sqrt3 r (SmoothedParticles.cpp:74-80)
    if dim == 2:
        ref = _w(2, r, 0) + 6 * _w(2, r, 2 * r) + 6 * _w(2, r, 2 * math.sqrt(3) * r)
    else:
        ref = (_w(3, r, 0) + 12 * _w(3, r, 2 * r) + 6 * _w(3, r, 2 * math.sqrt(2) * r)
               + 24 * _w(3, r, 2 * math.sqrt(3) * r))
    assert O.packed_kernel_sum(dim, r) == pytest.approx(ref, rel=1e-13)
