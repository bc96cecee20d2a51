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


def test_bf16_rounding_and_inference_forward_against_torch():
    """The reference's BF16 rounding (nearest even) and its inference forward,
    against torch's bfloat16 casts with float64 products: an independent
    implementation of the same rounding and the same network."""
    import torch
    rng = np.random.default_rng(31)
    edge = np.array([1.0, 1.00390625, 1.005859375, 1.001953125, 1.0009765625, -3.14159, 6e-39, 1e38, 0.0, -0.0],
                    dtype=np.float32)
    vals = np.concatenate([edge, rng.normal(0, 3, 10000).astype(np.float32)])
    want = torch.from_numpy(vals).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(M.bf16(vals), want)

    theta = rng.normal(0, 0.2, M.PARAMS).astype(np.float32)
    x = rng.normal(0, 1, (256, M.IN)).astype(np.float32)
    p = {k: torch.from_numpy(np.ascontiguousarray(v)).to(torch.bfloat16).to(torch.float64)
         for k, v in M.unpack(theta).items()}

    def rb(t):
        return t.to(torch.float32).to(torch.bfloat16).to(torch.float64)

    def gelu(z):
        return 0.5 * z * (1.0 + torch.erf(z / np.sqrt(2.0)))

    h0 = rb(torch.from_numpy(x).to(torch.float64))
    h1 = rb(gelu(h0 @ p["W1"].T + p["b1"]))
    h2 = rb(gelu(h1 @ p["W2"].T + p["b2"]))
    y = (h2 @ p["W3"].T + p["b3"]).numpy()
    got = M.forward_infer(theta, x)
    # float32 against float64 accumulation; a rounding flip of one activation moves
    # an output by ~1e-4 at these weights
    err = np.abs(got - y)
    assert np.mean(err < 1e-5) > 0.9 and err.max() < 2e-3
