"""CPU oracle of the sizing model. TEST INFRASTRUCTURE ONLY: used by tests/ to check the GPU
forward, backward and Adam kernels; the product never imports it.

The network (no reference counterpart; the sizing model is new in this build): 16 inputs ->
32 ReLU -> 32 ReLU -> 1, parameters flat in the order W1[32][16], b1[32], W2[32][32], b2[32],
w3[32], b3; loss = mean over the batch of (y - t)^2. Restated from the textbook definitions of a
multilayer perceptron, its backpropagation and Adam (Kingma & Ba 2015), in float64 unless asked
for float32. Pinned in tests by central finite differences of the loss (float64)."""
from __future__ import annotations

import numpy as np

N_IN, N_H = 16, 32
N_PARAMS = N_H * N_IN + N_H + N_H * N_H + N_H + N_H + 1


def unpack(p: np.ndarray):
    o = 0
    W1 = p[o:o + N_H * N_IN].reshape(N_H, N_IN)
    o += N_H * N_IN
    b1 = p[o:o + N_H]
    o += N_H
    W2 = p[o:o + N_H * N_H].reshape(N_H, N_H)
    o += N_H * N_H
    b2 = p[o:o + N_H]
    o += N_H
    w3 = p[o:o + N_H]
    o += N_H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def forward(p: np.ndarray, x: np.ndarray):
    """Returns (y, h1, h2, z1, z2) for inputs x (B x 16)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    z1 = x @ W1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ w3 + b3
    return y, h1, h2, z1, z2


def loss(p: np.ndarray, x: np.ndarray, t: np.ndarray) -> float:
    y = forward(p, x)[0]
    return float(np.mean((y - t) ** 2))


def grad(p: np.ndarray, x: np.ndarray, t: np.ndarray) -> np.ndarray:
    """Analytic gradient of the mean squared error with respect to the flat parameters."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    y, h1, h2, z1, z2 = forward(p, x)
    B = len(t)
    d3 = 2 * (y - t) / B
    d2 = np.outer(d3, w3) * (z2 > 0)
    d1 = (d2 @ W2) * (z1 > 0)
    g = np.concatenate([(d1.T @ x).ravel(), d1.sum(0), (d2.T @ h1).ravel(), d2.sum(0), h2.T @ d3, [d3.sum()]])
    return g


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step with bias correction; returns (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
