"""FP64 numpy restatement of the scattering surrogate (layout of mt_mlp.h):
forward pass, mean-squared loss and back-propagated gradient.  The checker for
the device kernels mlp_fwdbwd/mlp_grad; itself pinned by central finite
differences in tests/test_surrogate.py."""
from __future__ import annotations

import numpy as np

H = 64
OFF_W1, OFF_B1, OFF_W2 = 0, 128, 192
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
N_PARAMS = OFF_B3 + 1


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(H, 2), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def forward(p, X):
    """X: (B, 2) normalised inputs -> y (B,)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    h1 = np.tanh(X @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ W3 + b3


def loss_grad(p, X, T):
    """Mean squared error over the batch and its gradient (flat, N_PARAMS)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    B = len(X)
    h1 = np.tanh(X @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    y = h2 @ W3 + b3
    e = y - T
    dy = 2.0 * e / B
    d2 = (dy[:, None] * W3[None, :]) * (1.0 - h2 * h2)
    d1 = (d2 @ W2) * (1.0 - h1 * h1)
    g = np.zeros(N_PARAMS)
    g[OFF_W1:OFF_B1] = (d1.T @ X).ravel()
    g[OFF_B1:OFF_W2] = d1.sum(0)
    g[OFF_W2:OFF_B2] = (d2.T @ h1).ravel()
    g[OFF_B2:OFF_W3] = d2.sum(0)
    g[OFF_W3:OFF_B3] = h2.T @ dy
    g[OFF_B3] = dy.sum()
    return float(np.mean(e * e)), g


def normalise(eps, b, eps_min, eps_max, b_max):
    """Inputs as the device normalises them (float32 constants)."""
    l0, l1 = np.log10(eps_min), np.log10(eps_max)
    e_mid, e_inv = np.float32(0.5 * (l0 + l1)), np.float32(2.0 / (l1 - l0))
    b_mid, b_inv = np.float32(0.5 * b_max), np.float32(2.0 / b_max)
    x0 = ((np.log10(eps) - np.float64(e_mid)) * np.float64(e_inv)).astype(np.float32)
    x1 = ((b - np.float64(b_mid)) * np.float64(b_inv)).astype(np.float32)
    return np.stack([x0, x1], 1).astype(np.float64)


Y_OFF, Y_SCALE = -12.0, 12.0


def target(ln_s2):
    return ((np.asarray(ln_s2) - Y_OFF) / Y_SCALE).astype(np.float32).astype(np.float64)


def s2_of(p, eps, b, eps_min, eps_max, b_max):
    y = forward(p, normalise(eps, b, eps_min, eps_max, b_max))
    return np.minimum(np.exp(y * Y_SCALE + Y_OFF), 1.0)
