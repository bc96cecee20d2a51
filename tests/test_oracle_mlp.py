"""Pins the numpy restatement of the learned score's MLP (oracle/mlp_ref.py):
gradient by central finite differences in float64, GELU by its series at 0 and
its limits, BF16 rounding by hand-worked values, Adam by its first step."""
from __future__ import annotations

import math

import numpy as np

from oracle import mlp_ref as M


def test_parameter_count_matches_the_abi():
    from remobench_amd import lib as L
    assert M.PARAMS == 5186 == 64 * 13 + 64 + 64 * 64 + 64 + 2 * 64 + 2
    assert L.MLP_PARAMS == M.PARAMS


def test_bf16_rounding_by_hand():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e38], dtype=np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0, -2.5, 2.9996e38], dtype=np.float32)
    got = M.bf16(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even (1); 1 + 3*2^-8 ties up to 1 + 2^-6
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == np.float32(1.0 + 2 ** -6) and got[3] == 1.0
    assert got[4] == -2.5
    assert (got.view(np.uint32) & 0xFFFF).max() == 0
    np.testing.assert_allclose(got[5], want[5], rtol=4e-3)


def test_gelu_limits_and_slope():
    z = np.array([-30.0, 0.0, 30.0])
    np.testing.assert_allclose(M.gelu(z), [0.0, 0.0, 30.0], atol=1e-12)
    np.testing.assert_allclose(M.gelu_grad(np.array([0.0, 30.0, -30.0])), [0.5, 1.0, 0.0], atol=1e-12)
    # derivative by finite differences
    z = np.linspace(-4, 4, 33)
    h = 1e-6
    np.testing.assert_allclose((M.gelu(z + h) - M.gelu(z - h)) / (2 * h), M.gelu_grad(z), atol=1e-8)


def test_gradient_against_finite_differences():
    rng = np.random.default_rng(0)
    theta = rng.normal(0, 0.3, M.PARAMS)
    x = rng.normal(0, 1, (7, M.IN))
    t = rng.normal(0, 1, (7, M.OUT))
    loss, g = M.loss_and_grad(theta, x, t, dtype=np.float64)
    idx = rng.choice(M.PARAMS, 200, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros(M.PARAMS)
        e[i] = h
        lp, _ = M.loss_and_grad(theta + e, x, t, dtype=np.float64)
        lm, _ = M.loss_and_grad(theta - e, x, t, dtype=np.float64)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_adam_first_step_is_lr_times_sign():
    # with bias correction the first step moves every parameter by lr * g / (|g| + eps)
    g = np.array([0.3, -2.0, 1e-3])
    theta, m, v = M.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, 1e-3)
    np.testing.assert_allclose(theta, -1e-3 * np.sign(g), rtol=1e-4)
