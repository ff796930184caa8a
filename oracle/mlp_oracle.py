"""FP64 CPU restatement of the learned initialiser's network (test infrastructure only).

Same function as the GPU kernels of bia_3d_gpu_amd/csrc/bia_learn.hip: inputs the first three
columns of a row, two GELU (tanh form) layers of 64, linear output; loss = mean((y - t)^2).
Parameters: W1[64][3] b1[64] W2[64][64] b2[64] w3[64] b3 (4481 floats).  There is no reference
counterpart; the gradient is pinned by central finite differences in tests/test_mlp.py.
"""
from __future__ import annotations

import numpy as np

H = 64
NPARAM = 3 * H + H + H * H + H + H + 1
C = np.sqrt(2.0 / np.pi)


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    W1 = w[0:192].reshape(H, 3)
    b1 = w[192:256]
    W2 = w[256:4352].reshape(H, H)
    b2 = w[4352:4416]
    w3 = w[4416:4480]
    b3 = w[4480]
    return W1, b1, W2, b2, w3, b3


def gelu(z):
    return 0.5 * z * (1 + np.tanh(C * (z + 0.044715 * z ** 3)))


def gelu_d(z):
    th = np.tanh(C * (z + 0.044715 * z ** 3))
    return 0.5 * (1 + th) + 0.5 * z * (1 - th * th) * C * (1 + 3 * 0.044715 * z * z)


def inputs(rows):
    rows = np.asarray(rows, dtype=np.float64)
    return np.ascontiguousarray(rows[:, :3])


def forward(w, rows, cache=False):
    W1, b1, W2, b2, w3, b3 = unpack(w)
    X = inputs(rows)
    Z1 = X @ W1.T + b1
    H1 = gelu(Z1)
    Z2 = H1 @ W2.T + b2
    H2 = gelu(Z2)
    y = H2 @ w3 + b3
    return (y, (X, Z1, H1, Z2, H2)) if cache else y


def loss_and_grad(w, rows):
    rows = np.asarray(rows, dtype=np.float64)
    W1, b1, W2, b2, w3, b3 = unpack(w)
    y, (X, Z1, H1, Z2, H2) = forward(w, rows, cache=True)
    t = rows[:, 3]
    B = len(rows)
    loss = np.mean((y - t) ** 2)
    dy = 2 * (y - t) / B
    g = np.zeros(NPARAM)
    g[4416:4480] = H2.T @ dy
    g[4480] = dy.sum()
    dZ2 = np.outer(dy, w3) * gelu_d(Z2)
    g[256:4352] = (dZ2.T @ H1).reshape(-1)
    g[4352:4416] = dZ2.sum(0)
    dZ1 = (dZ2 @ W2) * gelu_d(Z1)
    g[0:192] = (dZ1.T @ X).reshape(-1)
    g[192:256] = dZ1.sum(0)
    return loss, g
