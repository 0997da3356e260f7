"""Oracle (test infrastructure only) for the learned SHAKE warm start: the
multilayer perceptron's forward pass, loss and gradients in float64 numpy.

The warm start has no counterpart in the reference (it attaches where
`shake_cluster` starts, src/gpu/kernels/shake.wgsl:130-146, and to the
oracle's multiplier accumulation, src/reference/constraints.ts:147-184).  This
restates the network the device trains and runs, with the device's rounding
points: every matrix operand is rounded to bfloat16 (nearest, ties to even)
and every product is accumulated in higher precision.

    H1 = tanh(bf(X) bf(W1) + b1)          X: [n, 32], W1: [32, 64]
    H2 = tanh(bf(H1) bf(W2) + b2)         W2: [64, 64]
    Y  = bf(H2) bf(W3) + b3               W3: [64, 16], outputs 0..3 used
    loss = sum(mask (Y - T)^2) / (4 n)      over the 4 used outputs, where
                                            mask marks the constraints present

with the gradients of that loss through the same rounded operands.
`rounding=False` drops the bfloat16 rounding, which is how the tests pin the
gradient formulas themselves against central differences.
"""
from __future__ import annotations

import numpy as np

F, H, O, USED = 32, 64, 16, 4
# feature slot of constraint c's "present" flag: features 4c..4c+3 are
# (scaled violation, scaled projection, reduced-mass factor, present)
PRESENT = (3, 7, 11, 15)


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as
    float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def mask_of(X):
    m = np.zeros((X.shape[0], O))
    for c, k in enumerate(PRESENT):
        m[:, c] = X[:, k] > 0.5
    return m


def _identity(x):
    return np.asarray(x, dtype=np.float64)


def forward(X, W1, b1, W2, b2, W3, b3, rounding=True):
    r = bf16 if rounding else _identity
    Xb = r(X)
    A1 = Xb @ r(W1) + b1
    H1 = np.tanh(A1)
    A2 = r(H1) @ r(W2) + b2
    H2 = np.tanh(A2)
    Y = r(H2) @ r(W3) + b3
    return Y, (Xb, H1, H2)


def loss_and_gradients(X, T, W1, b1, W2, b2, W3, b3, rounding=True):
    """(loss, dict of gradients) for targets T [n, 4]."""
    r = bf16 if rounding else _identity
    bf = r
    Y, (Xb, H1, H2) = forward(X, W1, b1, W2, b2, W3, b3, rounding)
    M = mask_of(X)
    Tp = np.zeros_like(Y)
    Tp[:, :USED] = T
    count = 4.0 * X.shape[0]
    diff = (Y - Tp) * M
    loss = float(np.sum(diff * diff) / count)
    dY = 2.0 * diff / count
    g = {}
    g["W3"] = bf(H2).T @ bf(dY)
    g["b3"] = dY.sum(axis=0)
    dA2 = (bf(dY) @ bf(W3).T) * (1.0 - H2 * H2)
    g["W2"] = bf(H1).T @ bf(dA2)
    g["b2"] = dA2.sum(axis=0)
    dA1 = (bf(dA2) @ bf(W2).T) * (1.0 - H1 * H1)
    g["W1"] = Xb.T @ bf(dA1)
    g["b1"] = dA1.sum(axis=0)
    return loss, g


# parameter vector layout shared with the device: W1 | b1 | W2 | b2 | W3 | b3
LAYOUT = (("W1", (F, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, O)), ("b3", (O,)))
NPARAMS = sum(int(np.prod(shape)) for _, shape in LAYOUT)


def unpack(theta):
    out, at = {}, 0
    for name, shape in LAYOUT:
        size = int(np.prod(shape))
        out[name] = np.asarray(theta[at:at + size], dtype=np.float64).reshape(shape)
        at += size
    return out


def pack(params):
    return np.concatenate([np.asarray(params[name], dtype=np.float64).reshape(-1) for name, _ in LAYOUT])
