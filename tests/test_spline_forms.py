"""The rearranged forms of the cubic spline that the FP32 kernels evaluate
(SmoothedParticles.h:60-72), checked against the oracle's piecewise spline
and derivative (themselves pinned by closed forms in test_oracle_kernel.py):

- W = c0 (2 a^3 - b^3), a = max(1 - x, 0), b = max(1 - 2x, 0): one
  expression for both pieces and zero support (density walk, virtual sums);
- S = -W'(r) / r = c1 (a^2 - b^2) / r, c1 = 6 c0 / h (gradient, virtual sums);
- in units of H (the gradient): w = (r/H)^2, u = 1/sqrt(w), a = 1 - w u and
  b = 2a - 1, both clamped to [0, 1], with S d = c1 (a^2 - b^2) u (d / H);
- the density's split cube sums 2 sum(a^3) - sum(c^3) in FP32 against the
  FP64 sum of the terms (no cancellation: 2a^3 >= c^3 term by term)."""
import math

import numpy as np
import pytest

import oracle_py as O


def _ab(x):
    return np.maximum(1 - x, 0.0), np.maximum(1 - 2 * x, 0.0)


@pytest.mark.parametrize("dim", [2, 3])
def test_two_cube_form_is_the_spline(dim):
    r = 0.01
    h = 4 * r
    c0 = O.kernel(dim, r, 0.0)
    for d in np.linspace(0.0, 1.3 * h, 131):
        a, b = _ab(d / h)
        w = c0 * (2 * a ** 3 - b ** 3)
        assert w == pytest.approx(O.kernel(dim, r, d), rel=1e-12, abs=1e-14 * c0)


@pytest.mark.parametrize("dim", [2, 3])
def test_gradient_factor_form(dim):
    r = 0.01
    h = 4 * r
    c1 = 6 * O.kernel(dim, r, 0.0) / h
    for d in np.linspace(1e-3 * h, 1.3 * h, 131):
        a, b = _ab(d / h)
        s = c1 * (a * a - b * b) / d
        ref = -O.kernel_d1(dim, r, d) / d
        assert s == pytest.approx(ref, rel=1e-12, abs=1e-14 * c1 / h)


def test_h_unit_form_of_the_gradient():
    # the virtual particles' kernel has support H = 4 s (radius s)
    s_ = 0.01
    H = 4 * s_
    c1 = 6 * O.kernel(3, s_, 0.0) / H
    rng = np.random.default_rng(7)
    for _ in range(400):
        d = rng.uniform(-1.6 * H, 1.6 * H, size=3)
        r = float(np.linalg.norm(d))
        ds = d / H
        w = float(ds @ ds)
        u = 1.0 / math.sqrt(w)
        a = min(max(1 - w * u, 0.0), 1.0)
        b = min(max(2 * a - 1, 0.0), 1.0)
        a0, b0 = _ab(r / H)
        assert a == pytest.approx(a0, abs=1e-14) and b == pytest.approx(b0, abs=1e-14)
        sd = c1 * (a * a - b * b) * u * ds
        ref = -O.kernel_d1(3, s_, r) / r * d
        assert np.allclose(sd, ref, rtol=1e-11, atol=1e-13 * c1)


def test_split_cube_sums_in_fp32():
    rng = np.random.default_rng(11)
    for _ in range(50):
        x = rng.uniform(0.0, 1.2, size=rng.integers(1, 220)).astype(np.float32)
        a = np.maximum(np.float32(1) - x, np.float32(0))
        c = np.maximum(np.float32(1) - np.float32(2) * x, np.float32(0))
        suma = np.float32(0)
        sumc = np.float32(0)
        for ai, ci in zip(a, c):
            suma = np.float32(suma + ai * ai * ai)
            sumc = np.float32(sumc + ci * ci * ci)
        got = np.float32(np.float32(2) * suma - sumc)
        a64, c64 = _ab(x.astype(np.float64))
        ref = float(np.sum(2 * a64 ** 3 - c64 ** 3))
        assert float(got) == pytest.approx(ref, rel=2e-6, abs=1e-6)
