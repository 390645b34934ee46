"""Pins the FP64 MLP restatement (oracle/mlp_oracle.py) without a GPU: its backward against
central finite differences of its own exact forward, and the BF16 rounding helper against
hand-checked values."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp_oracle as M


def test_bf16_rounding_nearest_even():
    # 1 + 2^-8 is halfway between bf16 neighbours 1 and 1 + 2^-7: ties to even -> 1
    assert M.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0
    # 1 + 3 * 2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: ties to even -> 1 + 2^-6
    assert M.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6
    assert M.bf16(np.float32(-2.5)) == -2.5
    assert M.bf16(np.float32(0.1)) == pytest.approx(0.1, rel=4e-3)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = M.init_params(1) * 0.5
    p[M.OFF_B1:M.OFF_W2] = rng.standard_normal(64) * 0.1
    p[M.OFF_B2:M.OFF_W3] = rng.standard_normal(64) * 0.1
    X = rng.standard_normal((40, 8))
    T = rng.standard_normal(40)
    loss, g = M.loss_and_grad(p, X, T, bf16_ops=False)
    idx = np.concatenate([rng.choice(np.arange(M.OFF_W1, M.OFF_B1), 8), rng.choice(np.arange(M.OFF_B1, M.OFF_W2), 4),
                          rng.choice(np.arange(M.OFF_W2, M.OFF_B2), 12), rng.choice(np.arange(M.OFF_B2, M.OFF_W3), 4),
                          rng.choice(np.arange(M.OFF_W3, M.OFF_B3), 6), [M.OFF_B3]])
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (M.loss_and_grad(pp, X, T, False)[0] - M.loss_and_grad(pm, X, T, False)[0]) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-5, abs=1e-8), k


def test_forward_shapes_and_relu():
    p = np.zeros(M.N_PARAM)
    p[M.OFF_B3] = 0.25
    y, _ = M.forward(p, np.ones((5, 8)))
    assert np.array_equal(y, np.full(5, 0.25))
