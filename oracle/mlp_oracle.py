"""FP64 restatement of the warm-start MLP (6 -> 64 -> 64 -> 1, ReLU) -- TEST INFRASTRUCTURE ONLY.

Parameter layout (flat, 4673): W1[6][64], b1[64], W2[64][64], b2[64], w3[64], b3.
`bf16=True` rounds operands to BF16 (round to nearest even) at the points where the gfx950
kernels do: the inputs x, W1, W2, w3, the hidden activation h1, and the backward signals dH2,
dH1 that enter the weight-gradient products; everything else is FP64.  The backward is
pinned by finite differences of the exact (bf16=False) forward in tests/test_mlp_oracle.py.
"""

from __future__ import annotations

import numpy as np

N_IN, N_H = 6, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_IN * N_H
OFF_W2 = OFF_B1 + N_H
OFF_B2 = OFF_W2 + N_H * N_H
OFF_W3 = OFF_B2 + N_H
OFF_B3 = OFF_W3 + N_H
N_PARAM = OFF_B3 + 1


def bf16(x) -> np.ndarray:
    """Round to BF16 (nearest even) and return as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    return np.where(np.isfinite(a), out, a.astype(np.float64))


def unpack(p: np.ndarray):
    p = np.asarray(p, dtype=np.float64)
    W1 = p[OFF_W1:OFF_B1].reshape(N_IN, N_H)
    b1 = p[OFF_B1:OFF_W2]
    W2 = p[OFF_W2:OFF_B2].reshape(N_H, N_H)
    b2 = p[OFF_B2:OFF_W3]
    w3 = p[OFF_W3:OFF_B3]
    b3 = p[OFF_B3]
    return W1, b1, W2, b2, w3, b3


def forward(p, X, bf16_ops: bool = True):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    x = np.asarray(X, dtype=np.float64)[:, :N_IN]
    q = bf16 if bf16_ops else (lambda v: np.asarray(v, dtype=np.float64))
    xb, W1b, W2b, w3b = q(x), q(W1), q(W2), q(w3)
    a1 = xb @ W1b + b1
    h1 = np.maximum(a1, 0.0)
    h1b = q(h1)
    a2 = h1b @ W2b + b2
    h2 = np.maximum(a2, 0.0)
    y = h2 @ w3b + b3
    return y, dict(xb=xb, W1b=W1b, W2b=W2b, w3b=w3b, a1=a1, h1=h1, h1b=h1b, a2=a2, h2=h2)


def loss_and_grad(p, X, T, bf16_ops: bool = True):
    """Mean squared error and its gradient (flat, FP64)."""
    y, c = forward(p, X, bf16_ops)
    q = bf16 if bf16_ops else (lambda v: np.asarray(v, dtype=np.float64))
    n = y.shape[0]
    e = y - np.asarray(T, dtype=np.float64)
    loss = float(np.mean(e * e))
    dy = 2.0 * e / n
    g = np.zeros(N_PARAM)
    g[OFF_W3:OFF_B3] = c["h2"].T @ dy
    g[OFF_B3] = dy.sum()
    dH2 = np.outer(dy, c["w3b"]) * (c["a2"] > 0)
    g[OFF_B2:OFF_W3] = dH2.sum(axis=0)
    dH2b = q(dH2)
    g[OFF_W2:OFF_B2] = (c["h1b"].T @ dH2b).reshape(-1)
    dH1 = (dH2b @ c["W2b"].T) * (c["h1b"] > 0)
    g[OFF_B1:OFF_W2] = dH1.sum(axis=0)
    g[OFF_W1:OFF_B1] = (c["xb"].T @ q(dH1)).reshape(-1)
    return loss, g


def init_params(seed: int = 0) -> np.ndarray:
    """He-style initialisation (the GPU trainer uses the same)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAM)
    p[OFF_W1:OFF_B1] = rng.standard_normal(N_IN * N_H) * np.sqrt(2.0 / N_IN)
    p[OFF_W2:OFF_B2] = rng.standard_normal(N_H * N_H) * np.sqrt(2.0 / N_H)
    p[OFF_W3:OFF_B3] = rng.standard_normal(N_H) * np.sqrt(1.0 / N_H)
    return p
