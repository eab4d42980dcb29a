"""FP64 numpy restatement of the scattering surrogate (layout of mt_mlp.h):
inputs, forward pass, mean-squared loss and back-propagated gradient.  The
checker for the device kernels mlp_fwdbwd/mlp_grad; itself pinned by central
finite differences in tests/test_surrogate.py."""
from __future__ import annotations

import numpy as np

H, NIN = 64, 3
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * NIN
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
N_PARAMS = OFF_B3 + 1


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(H, NIN), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def forward(p, X):
    """X: (B, 3) normalised inputs -> y (B,)."""
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


def inputs(norm, eps, b):
    """The three network inputs in float32, as mlp_inputs forms them."""
    f = np.float32
    e = np.asarray(eps, f)
    bb = np.asarray(b, f)
    x0 = (np.log10(e) - f(norm["e_mid"])) * f(norm["e_inv_half"])
    x1 = (np.log10(np.maximum(bb, f(norm["b_floor"]))) - f(norm["lb_mid"])) * f(norm["lb_inv_half"])
    x2 = (bb - f(norm["b_mid"])) * f(norm["b_inv_half"])
    return np.stack([x0, x1, x2], 1).astype(np.float64)


def target(norm, ln_s2):
    return ((np.asarray(ln_s2) - norm["y_off"]) / norm["y_scale"]).astype(np.float32).astype(np.float64)


def s2_of(p, norm, eps, b):
    y = forward(p, inputs(norm, eps, b))
    return np.minimum(np.exp(y * norm["y_scale"] + norm["y_off"]), 1.0)
