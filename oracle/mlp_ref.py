"""numpy restatement of the learned score's MLP (remobench_amd/csrc/mlp.hip).

TEST INFRASTRUCTURE ONLY: used by tests/, never by the product path.

13 -> 64 -> 64 -> 2 with exact GELU, parameters in the C ABI's order
(W1[64][13], b1[64], W2[64][64], b2[64], W3[2][64], b3[2]); the training forward
and backward passes run on BF16-rounded parameters (round to nearest even) with
float32 (or float64, for finite-difference checks) activations, the inference
forward rounds each layer's inputs to BF16 as well; the loss is the
mean squared error over n x 2 outputs; Adam with bias correction. There is no
reference implementation to follow (the reference's plans/04_NeuralPass.md is
empty); the gradient is pinned by finite differences in tests/test_oracle_mlp.py.
"""
from __future__ import annotations

import math

import numpy as np

IN, HID, OUT = 13, 64, 2
SIZES = [("W1", (HID, IN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (OUT, HID)), ("b3", (OUT,))]
PARAMS = sum(int(np.prod(s)) for _, s in SIZES)

_erf = np.vectorize(math.erf, otypes=[np.float64])


def unpack(theta):
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = np.asarray(theta[o:o + k]).reshape(shape)
        o += k
    return out


def bf16(x) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(np.asarray(z, dtype=np.float64) / math.sqrt(2.0))).astype(z.dtype)


def gelu_grad(z):
    z64 = np.asarray(z, dtype=np.float64)
    cdf = 0.5 * (1.0 + _erf(z64 / math.sqrt(2.0)))
    pdf = np.exp(-0.5 * z64 * z64) / math.sqrt(2.0 * math.pi)
    return (cdf + z64 * pdf).astype(z.dtype)


def forward(theta, x, dtype=np.float32, cache=False):
    p = {k: v.astype(dtype) for k, v in unpack(theta).items()}
    x = np.asarray(x, dtype=dtype)
    z1 = x @ p["W1"].T + p["b1"]
    h1 = gelu(z1)
    z2 = h1 @ p["W2"].T + p["b2"]
    h2 = gelu(z2)
    y = h2 @ p["W3"].T + p["b3"]
    if cache:
        return y, (x, z1, h1, z2, h2)
    return y


def forward_infer(theta, x):
    """The inference kernels' forward: BF16 parameters (theta as stored), every
    layer's inputs rounded to BF16 for the matrix cores, float32 accumulation and
    bias, exact GELU."""
    p = {k: v.astype(np.float32) for k, v in unpack(bf16(theta)).items()}
    h0 = bf16(np.asarray(x, dtype=np.float32))
    h1 = bf16(gelu(h0 @ p["W1"].T + p["b1"]))
    h2 = bf16(gelu(h1 @ p["W2"].T + p["b2"]))
    return h2 @ p["W3"].T + p["b3"]


def loss_and_grad(theta, x, t, dtype=np.float32):
    """Loss = mean over n x 2 of (y - t)^2, and d loss / d theta (flat)."""
    p = {k: v.astype(dtype) for k, v in unpack(theta).items()}
    y, (x, z1, h1, z2, h2) = forward(theta, x, dtype, cache=True)
    n = x.shape[0]
    d = y - np.asarray(t, dtype=dtype)
    loss = float((d.astype(np.float64) ** 2).sum() / (n * OUT))
    dy = 2.0 * d / (n * OUT)
    gW3 = dy.T @ h2
    gb3 = dy.sum(0)
    dz2 = (dy @ p["W3"]) * gelu_grad(z2)
    gW2 = dz2.T @ h1
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ p["W2"]) * gelu_grad(z1)
    gW1 = dz1.T @ x
    gb1 = dz1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g


def adam(theta, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam update (step counts from 1). Returns theta, m, v."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v
