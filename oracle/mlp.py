"""FP64 numpy restatement of the Newton-start network (test infrastructure).

The network has no counterpart in the reference: it seeds the Newton
intersection of surface.cpp:22-49 / sum.cpp:47-56 with a learned start.
This module is the checker of the GPU training kernels: forward pass, mean
squared error, and the analytic gradient of that loss, all in FP64.

Parameter vector layout (the library's): W1[64][6], b1[64], W2[64][64],
b2[64], w3[64], b3.  Inputs are normalised as (x - in_mean) * in_inv_std.

Pinned by: central finite differences of ``loss`` (tests/test_mlp_oracle.py),
which fix the backward pass independently of how it is written, and by a
one-neuron closed form of the forward pass.
"""
import numpy as np

N_IN = 6
N_HID = 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_HID * N_IN
OFF_W2 = OFF_B1 + N_HID
OFF_B2 = OFF_W2 + N_HID * N_HID
OFF_W3 = OFF_B2 + N_HID
OFF_B3 = OFF_W3 + N_HID
N_PARAMS = OFF_B3 + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    W1 = p[OFF_W1:OFF_B1].reshape(N_HID, N_IN)
    b1 = p[OFF_B1:OFF_W2]
    W2 = p[OFF_W2:OFF_B2].reshape(N_HID, N_HID)
    b2 = p[OFF_B2:OFF_W3]
    w3 = p[OFF_W3:OFF_B3]
    b3 = p[OFF_B3]
    return W1, b1, W2, b2, w3, b3


def normalise(features, in_norm):
    in_norm = np.asarray(in_norm, dtype=np.float64)
    return (np.asarray(features, dtype=np.float64) - in_norm[:N_IN]) * in_norm[N_IN:2 * N_IN]


def forward(p, features, in_norm):
    """Network output in normalised units, shape [n]."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    x = normalise(features, in_norm)
    h1 = np.maximum(x @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ w3 + b3


def loss(p, features, targets, in_norm):
    y = forward(p, features, in_norm)
    return float(np.mean((y - np.asarray(targets, dtype=np.float64)) ** 2))


def loss_grad(p, features, targets, in_norm):
    """(loss, dloss/dp) by reverse-mode differentiation."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    t = np.asarray(targets, dtype=np.float64)
    n = len(t)
    x = normalise(features, in_norm)
    a1 = x @ W1.T + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0.0)
    y = h2 @ w3 + b3
    err = y - t
    dy = 2.0 * err / n
    g = np.zeros(N_PARAMS)
    g[OFF_B3] = dy.sum()
    g[OFF_W3:OFF_B3] = h2.T @ dy
    g2 = np.outer(dy, w3) * (a2 > 0)
    g[OFF_B2:OFF_W3] = g2.sum(axis=0)
    g[OFF_W2:OFF_B2] = (g2.T @ h1).ravel()
    g1 = (g2 @ W2) * (a1 > 0)
    g[OFF_B1:OFF_W2] = g1.sum(axis=0)
    g[OFF_W1:OFF_B1] = (g1.T @ x).ravel()
    return float(np.mean(err ** 2)), g


def adam(p, g, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1); returns (p, m, v)."""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** step)
    vh = v / (1 - beta2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
