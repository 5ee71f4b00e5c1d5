"""FP64 CPU oracle of the per-face flux-correction network (test
infrastructure only: imported by tests/, __graft_entry__.smoke() and the
checker legs of bench.py; never by the product path).

The network has no counterpart in the FALL3D reference. It is the learned
part that corrects the coarse-grid KT face fluxes: a multilayer perceptron
32 -> 64 -> 64 -> 1 (31 features padded to 32), GELU activations (tanh
form), mean-squared-error loss and Adam updates. The GPU kernels compute the
same function with BF16 weights and activations and FP32 accumulation; this
module computes it in FP64 and is pinned by finite differences
(tests/test_mlp_oracle.py) and by closed forms of GELU.

Parameter layout (shared with the device code):
  W1 (32, 64), b1 (64,), W2 (64, 64), b2 (64,), w3 (64,), b3 ()
  z1 = x W1 + b1, h1 = gelu(z1), z2 = h1 W2 + b2, h2 = gelu(z2),
  y = h2 . w3 + b3.
"""
from __future__ import annotations

import math

import numpy as np

NIN, NH = 32, 64
SQRT_2_OVER_PI = math.sqrt(2.0 / math.pi)
GELU_C = 0.044715

# order and shapes of the flat parameter vector
PARAM_SHAPES = (("W1", (NIN, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("w3", (NH,)), ("b3", ()))
NPARAM = sum(int(np.prod(s)) for _, s in PARAM_SHAPES)


def gelu(x):
    """0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))."""
    return 0.5 * x * (1.0 + np.tanh(SQRT_2_OVER_PI * (x + GELU_C * x ** 3)))


def gelu_grad(x):
    u = SQRT_2_OVER_PI * (x + GELU_C * x ** 3)
    t = np.tanh(u)
    du = SQRT_2_OVER_PI * (1.0 + 3.0 * GELU_C * x * x)
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * du


def bf16_round(a):
    """Round to bfloat16 (nearest, ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = ((u + 0x7FFF + lsb) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def init_params(seed=0):
    """He-style initialisation (normal, std sqrt(2/fan_in)); biases zero."""
    rng = np.random.default_rng(seed)
    p = {
        "W1": rng.standard_normal((NIN, NH)) * math.sqrt(2.0 / NIN),
        "b1": np.zeros(NH),
        "W2": rng.standard_normal((NH, NH)) * math.sqrt(2.0 / NH),
        "b2": np.zeros(NH),
        "w3": rng.standard_normal(NH) * math.sqrt(1.0 / NH),
        "b3": np.zeros(()),
    }
    return p


def flatten(p):
    return np.concatenate([np.asarray(p[k], dtype=np.float64).ravel() for k, _ in PARAM_SHAPES])


def unflatten(v):
    out, o = {}, 0
    for k, s in PARAM_SHAPES:
        n = int(np.prod(s))
        out[k] = np.asarray(v[o:o + n], dtype=np.float64).reshape(s)
        o += n
    return out


def forward(p, x):
    """x (B, 32) -> y (B,), with the activations kept for backward."""
    z1 = x @ p["W1"] + p["b1"]
    h1 = gelu(z1)
    z2 = h1 @ p["W2"] + p["b2"]
    h2 = gelu(z2)
    y = h2 @ p["w3"] + p["b3"]
    return y, (x, z1, h1, z2, h2)


def loss_and_grad(p, x, t):
    """Mean squared error 1/B sum (y - t)^2 and its gradient (dict like p)."""
    y, (x, z1, h1, z2, h2) = forward(p, x)
    B = x.shape[0]
    r = y - t
    loss = float(np.mean(r * r))
    dy = 2.0 * r / B
    g = {"w3": h2.T @ dy, "b3": np.asarray(dy.sum())}
    dz2 = np.outer(dy, p["w3"]) * gelu_grad(z2)
    g["W2"] = h1.T @ dz2
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ p["W2"].T) * gelu_grad(z1)
    g["W1"] = x.T @ dz1
    g["b1"] = dz1.sum(0)
    return loss, g


def adam_step(v, g, m, s, step, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update of the flat vectors (Kingma & Ba, 2015, Algorithm 1);
    step counts from 1. Returns new (v, m, s)."""
    m = beta1 * m + (1.0 - beta1) * g
    s = beta2 * s + (1.0 - beta2) * g * g
    mh = m / (1.0 - beta1 ** step)
    sh = s / (1.0 - beta2 ** step)
    return v - lr * mh / (np.sqrt(sh) + eps), m, s
