"""Pins of the FP64 MLP oracle: analytic backprop vs central finite
differences, GELU and its derivative, bf16 rounding."""
import numpy as np

import mlp_oracle as M


def test_gelu_derivative():
    z = np.linspace(-4, 4, 81)
    e = 1e-6
    fd = (M.gelu(z + e) - M.gelu(z - e)) / (2 * e)
    np.testing.assert_allclose(M.gelu_d(z), fd, rtol=1e-7, atol=1e-9)
    assert M.gelu(0.0) == 0.0
    np.testing.assert_allclose(M.gelu(10.0), 10.0, rtol=1e-12)
    # the tanh form stays within 1e-3 of the exact erf GELU
    assert np.abs(M.gelu(z) - M.gelu_erf(z)).max() < 1e-3
    # the form used by the kernels: z * sigmoid(2u)
    u2 = 2 * M.C_GELU * (z + 0.044715 * z ** 3)
    np.testing.assert_allclose(M.gelu(z), z / (1 + np.exp(-u2)), rtol=1e-12, atol=1e-15)


def test_gradient_matches_finite_differences():
    p = M.random_params(3)
    rng = np.random.default_rng(4)
    X = rng.normal(size=(40, 8))
    t = rng.normal(size=40)
    loss, g = M.loss_and_grad(p, X, t)
    idx = np.concatenate([rng.choice(512, 20, replace=False),
                          512 + rng.choice(64, 8, replace=False),
                          576 + rng.choice(4096, 30, replace=False),
                          4672 + rng.choice(64, 8, replace=False),
                          4736 + rng.choice(64, 8, replace=False), [4800]])
    for i in idx:
        e = 1e-6 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += e
        pm[i] -= e
        fd = (M.loss_and_grad(pp, X, t)[0] - M.loss_and_grad(pm, X, t)[0]) / (2 * e)
        assert abs(g[i] - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(g).max()), i


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 1e-3], dtype=np.float32)
    r = M.bf16(x)
    assert r[0] == 1.0 and r[3] == -2.5
    assert r[1] == 1.0          # tie to even
    assert r[2] == 1.0 + 2 ** -7
    assert abs(r[4] - 1e-3) / 1e-3 < 2 ** -8
