"""FP64 restatement of the warm-start MLP (8 -> 64 -> 64 -> 1, tanh-form GELU).

TEST INFRASTRUCTURE ONLY (the checker of the hand-written MFMA kernels).
The network has no counterpart in the reference.  Parameter packing is the
one include/vps.h documents: W1[64x8], b1[64], W2[64x64] (out, in), b2[64],
W3[64], b3[1].  With ``bf16=True`` weights and the layer inputs are rounded
to bfloat16 (round to nearest even) where the GPU kernels round them.
"""
import numpy as np
from scipy.special import erf

N_IN, N_HID = 8, 64
N_PARAMS = N_HID * N_IN + N_HID + N_HID * N_HID + N_HID + N_HID + 1   # 4801


def bf16(a):
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + 512].reshape(64, 8); o += 512
    b1 = p[o:o + 64]; o += 64
    W2 = p[o:o + 4096].reshape(64, 64); o += 4096
    b2 = p[o:o + 64]; o += 64
    W3 = p[o:o + 64]; o += 64
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


C_GELU = np.sqrt(2 / np.pi)


def gelu(z):
    """GELU, tanh form: 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3)))."""
    return 0.5 * z * (1 + np.tanh(C_GELU * (z + 0.044715 * z ** 3)))


def gelu_d(z):
    t = np.tanh(C_GELU * (z + 0.044715 * z ** 3))
    return 0.5 * (1 + t) + 0.5 * z * (1 - t * t) * C_GELU * (1 + 3 * 0.044715 * z * z)


def gelu_erf(z):
    """The exact (erf) GELU, which the tanh form approximates to ~1e-3."""
    return 0.5 * z * (1 + erf(z / np.sqrt(2)))


def forward(p, X, bf=False):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    X = r(X)
    Z1 = X @ r(W1).T + b1
    H1 = r(gelu(Z1))
    Z2 = H1 @ r(W2).T + b2
    H2 = r(gelu(Z2))
    out = H2 @ r(W3) + b3
    return out, (X, Z1, H1, Z2, H2)


def loss_and_grad(p, X, t, bf=False):
    """Loss = mean (out - t)^2 and its gradient (packed like p)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    out, (Xr, Z1, H1, Z2, H2) = forward(p, X, bf)
    n = len(X)
    e = out - t
    loss = float(e @ e) / n
    dout = 2 * e / n
    gW3 = H2.T @ dout
    gb3 = dout.sum()
    dZ2 = np.outer(dout, r(W3)) * gelu_d(Z2)
    gW2 = dZ2.T @ H1
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2 @ r(W2)) * gelu_d(Z1)
    gW1 = dZ1.T @ Xr
    gb1 = dZ1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return loss, g


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS)
    W1, b1, W2, b2, W3, b3 = (slice(0, 512), slice(512, 576), slice(576, 4672),
                              slice(4672, 4736), slice(4736, 4800), slice(4800, 4801))
    p[W1] = rng.uniform(-1, 1, 512) * np.sqrt(6 / 72) * scale
    p[b1] = rng.normal(0, 0.1, 64)
    p[W2] = rng.uniform(-1, 1, 4096) * np.sqrt(6 / 128) * scale
    p[b2] = rng.normal(0, 0.1, 64)
    p[W3] = rng.normal(0, 0.2, 64)
    p[b3] = 0.05
    return p
