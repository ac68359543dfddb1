"""TEST INFRASTRUCTURE ONLY -- FP64 numpy MLP: forward, MSE backward and Adam, the checker of the
libnfx surrogate kernels.

Architecture and conventions match ``nofluidx3d_amd.surrogate.mlp``: y = W x + b per layer
(W [out x in]), exact GELU 0.5 z (1 + erf(z / sqrt 2)) on hidden layers, identity output, loss
sum_ij w_i (Y - T)^2 / (number of elements) with optional row weights w, Adam with bias correction.  Pinned in tests/test_oracle_mlp.py
against torch autograd in float64 and against central differences.
"""

import numpy as np
from scipy.special import erf

SQ2 = np.sqrt(2.0)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / SQ2))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / SQ2)) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(Ws, bs, X):
    """returns Y and the per-layer (input, pre-activation) cache"""
    cache = []
    h = X
    for l, (W, b) in enumerate(zip(Ws, bs)):
        z = h @ W.T + b
        cache.append((h, z))
        h = gelu(z) if l < len(Ws) - 1 else z
    return h, cache


def loss_and_grads(Ws, bs, X, T, w=None):
    """loss = sum_ij w_i (Y - T)^2 / Y.size (w = 1 when None) and its parameter gradients"""
    Y, cache = forward(Ws, bs, X)
    n = Y.size
    w = np.ones(len(Y)) if w is None else np.asarray(w, np.float64)
    loss = np.sum(w[:, None] * (Y - T) ** 2) / n
    dz = 2.0 * w[:, None] * (Y - T) / n
    gW, gb = [None] * len(Ws), [None] * len(Ws)
    for l in range(len(Ws) - 1, -1, -1):
        h, _ = cache[l]
        gW[l] = dz.T @ h
        gb[l] = dz.sum(0)
        if l > 0:
            dh = dz @ Ws[l]
            dz = dh * gelu_grad(cache[l - 1][1])
    return loss, gW, gb


def adam(params, grads, ms, vs, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """p -= lr / (1 - b1^t) * m / (sqrt(v) / sqrt(1 - b2^t) + eps) (Kingma & Ba, eps on sqrt(v_hat))"""
    lr_t = lr / (1 - b1**t)
    bc2 = 1.0 / np.sqrt(1 - b2**t)
    for p, g, m, v in zip(params, grads, ms, vs):
        m *= b1
        m += (1 - b1) * g
        v *= b2
        v += (1 - b2) * g * g
        p -= lr_t * m / (np.sqrt(v) * bc2 + eps)
