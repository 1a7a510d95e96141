"""Oracle, learned cost model -- TEST INFRASTRUCTURE ONLY.

FP64 numpy restatement of the learned planner's network, used by tests/ as
the checker of the HIP forward, backward and Adam kernels (mlp.hip) and of
the host driver (learned.c).  Nothing in the product imports this module.
The network has no counterpart in the reference (its planner is only the
analytic model of src/adatm_bench.c:108-142), so it is pinned by mathematics:
the gradient by central finite differences, Adam's first step by its closed
form p - lr * g / (|g| + eps), both in tests/test_cpu_learned.py.

Layout of the parameter vector (as adatm_mlp_get_params returns it):
W1 (24 x 64, row-major in x out), b1 (64), W2 (64 x 64), b2, W3 (64 x 64),
b3, W4 (64 x 1), b4 (1).  Hidden layers use ReLU; the output is linear;
the loss is the mean squared error over the batch.
"""
from __future__ import annotations

import numpy as np

LAYERS = ((24, 64), (64, 64), (64, 64), (64, 1))


def nparams() -> int:
    return sum(i * o + o for i, o in LAYERS)


def unpack(p: np.ndarray):
    out, off = [], 0
    for i, o in LAYERS:
        W = p[off:off + i * o].reshape(i, o)
        off += i * o
        b = p[off:off + o]
        off += o
        out.append((W, b))
    return out


def forward(p: np.ndarray, X: np.ndarray):
    """(predictions (B,), cache) in FP64."""
    H = np.asarray(X, dtype=np.float64)
    cache = []
    layers = unpack(np.asarray(p, dtype=np.float64))
    for li, (W, b) in enumerate(layers):
        Z = H @ W + b
        cache.append((H, Z))
        H = np.maximum(Z, 0.0) if li + 1 < len(layers) else Z
    return H[:, 0], cache


def loss_grad(p: np.ndarray, X: np.ndarray, t: np.ndarray):
    """mean((y - t)^2) and its gradient w.r.t. p, by backpropagation."""
    y, cache = forward(p, X)
    B = y.shape[0]
    d = y - np.asarray(t, dtype=np.float64)
    loss = float(np.mean(d * d))
    dZ = (2.0 / B) * d[:, None]
    layers = unpack(np.asarray(p, dtype=np.float64))
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        H, _ = cache[li]
        grads[li] = (H.T @ dZ, dZ.sum(axis=0))
        if li > 0:
            dH = dZ @ W.T
            dZ = dH * (cache[li - 1][1] > 0.0)
    g = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])
    return loss, g


def adam_step(p, g, m, v, step: int, lr: float, b1: float = 0.9, b2: float = 0.999,
              eps: float = 1e-8):
    """One Adam update (step counts from 1); returns (p, m, v)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** step)
    vh = v / (1.0 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def he_init(seed: int = 0) -> np.ndarray:
    """A random parameter vector for tests (not the library's initialiser)."""
    rng = np.random.default_rng(seed)
    parts = []
    for i, o in LAYERS:
        a = np.sqrt(6.0 / i)
        parts += [rng.uniform(-a, a, i * o), rng.uniform(-0.1, 0.1, o)]
    return np.concatenate(parts)
