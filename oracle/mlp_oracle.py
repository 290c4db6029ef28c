"""CPU oracle of the learned closure's network (FP64, numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ (and bench.py's CPU leg),
never by the product.  The network is the closure's MLP: 30 inputs
(padded to 32), two hidden layers of 64 GELU units, one output, trained
on the mean squared error.  The reference has no learned part, so there is
no reference file to follow; the network is fixed by its definition:

  z1 = x W1 + b1,  h1 = gelu(z1),  z2 = h1 W2 + b2,  h2 = gelu(z2),
  y  = h2 . w3 + b3,  L = mean((y - t)^2),  gelu(z) = z Phi(z).

`forward` / `grad` compute it in FP64.  With `bf16=True` they round to
bfloat16 (round to nearest even) exactly where the device kernels hold
bfloat16 operands -- the weights W1, W2, w3, the inputs x, the
activations h1, h2 and the back-propagated pre-activation gradients g_z1,
g_z2 -- so that the device's FP32 accumulation is the only remaining
difference.  Parameters are packed as in the C ABI: W1 [32][64] (row =
input), b1 [64], W2 [64][64] (row = input), b2 [64], w3 [64], b3 [1].
"""
from __future__ import annotations

import math

import numpy as np

N_IN, N_PAD, N_H = 30, 32, 64
SIZES = [("W1", (N_PAD, N_H)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)), ("w3", (N_H,)), ("b3", (1,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shp in SIZES:
        k = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + k], dtype=np.float64).reshape(shp)
        o += k
    return out


def pack(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SIZES])


def bf16(a) -> np.ndarray:
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


_erf = np.vectorize(math.erf)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def _pad(x):
    x = np.asarray(x, dtype=np.float64)
    out = np.zeros((x.shape[0], N_PAD))
    out[:, :x.shape[1]] = x
    return out


def forward(p: np.ndarray, x: np.ndarray, bf16_ops: bool = False):
    """Outputs y [n] and the intermediates of the backward pass."""
    q = unpack(p)
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, dtype=np.float64))
    X = r(_pad(x))
    W1, W2, w3 = r(q["W1"]), r(q["W2"]), r(q["w3"])
    z1 = X @ W1 + q["b1"]
    h1 = r(gelu(z1))
    z2 = h1 @ W2 + q["b2"]
    h2 = r(gelu(z2))
    y = h2 @ w3 + q["b3"][0]
    return y, dict(X=X, W1=W1, W2=W2, w3=w3, z1=z1, h1=h1, z2=z2, h2=h2)


def loss(p, x, t, bf16_ops=False) -> float:
    y, _ = forward(p, x, bf16_ops)
    return float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))


def grad(p: np.ndarray, x: np.ndarray, t: np.ndarray, bf16_ops: bool = False):
    """Gradient of the mean squared error over the samples, and the loss."""
    y, c = forward(p, x, bf16_ops)
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, dtype=np.float64))
    n = len(y)
    gy = 2.0 * (y - np.asarray(t, dtype=np.float64)) / n
    g = {}
    g["w3"] = c["h2"].T @ gy
    g["b3"] = np.array([gy.sum()])
    gz2 = r(np.outer(gy, c["w3"]) * gelu_grad(c["z2"]))
    g["W2"] = c["h1"].T @ gz2
    g["b2"] = gz2.sum(axis=0)
    gh1 = gz2 @ c["W2"].T
    gz1 = r(gh1 * gelu_grad(c["z1"]))
    g["W1"] = c["X"].T @ gz1
    g["b1"] = gz1.sum(axis=0)
    return pack(g), float(np.mean((y - t) ** 2))


def init_params(seed: int) -> np.ndarray:
    """He-style initial weights (normal, std sqrt(2/fan_in)), zero biases."""
    rng = np.random.default_rng(seed)
    d = {"W1": rng.normal(0, math.sqrt(2.0 / N_IN), (N_PAD, N_H)),
         "b1": np.zeros(N_H),
         "W2": rng.normal(0, math.sqrt(2.0 / N_H), (N_H, N_H)),
         "b2": np.zeros(N_H),
         "w3": rng.normal(0, math.sqrt(1.0 / N_H), N_H),
         "b3": np.zeros(1)}
    d["W1"][N_IN:, :] = 0.0
    return pack(d)


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015) in FP64; returns p, m, v."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
