"""CPU oracle of the return-map surrogate network (test infrastructure only).

Only tests, the smoke check and the benchmark's CPU baseline may use this
module; the product never imports it.  The network has no reference
counterpart (the reference solves its return map exactly,
physics/materials/ding.py:168-221); this restates the network the build
defines -- 7 inputs, 4 hidden layers of 64 SiLU units, 4 outputs -- in numpy:

* ``forward`` / ``gradient``: FP64 forward pass and the gradient of the mean
  squared error over the batch and the 4 outputs, by reverse-mode
  differentiation written out layer by layer (pinned by central finite
  differences in tests/test_oracle_mlp.py);
* ``infer_bf16``: the matrix-core inference -- inputs normalised in FP32,
  weights and every layer's input rounded to BF16 (round to nearest even),
  products accumulated in FP32 from the bias, SiLU in FP32 -- with the
  accumulation order left free (the hardware's differs), so agreement is to
  FP32 rounding of the sums plus the occasional BF16 rounding flip.

Parameter vector: W1[64][8] b1[64] W2[64][64] b2 W3[64][64] b3 W4[64][64] b4
W5[4][64] b5[4], row-major [out][in]; input 8 of W1 is unused.
"""
from __future__ import annotations

import numpy as np

SHAPES = (("W1", (64, 8)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (64, 64)), ("b3", (64,)),
          ("W4", (64, 64)), ("b4", (64,)), ("W5", (4, 64)), ("b5", (4,)))
COUNT = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(theta):
    out, off = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(theta[off:off + n]).reshape(shape)
        off += n
    return out


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(theta, x):
    """FP64 outputs (n, 4) and pre-activations [(n, 64)] * 4 for normalised inputs x (n, 8)."""
    p = unpack(np.asarray(theta, np.float64))
    x = np.asarray(x, np.float64).copy()
    x[:, 7:] = 0.0
    zs = []
    a = x
    for k in range(1, 5):
        z = a @ p[f"W{k}"].T + p[f"b{k}"]
        zs.append(z)
        a = _silu(z)
    return a @ p["W5"].T + p["b5"], zs


def loss(theta, x, y):
    out, _ = forward(theta, x)
    return float(np.mean((out - y) ** 2))


def gradient(theta, x, y):
    """Gradient of mean((f(x) - y)^2) over the batch and the 4 outputs, as a flat FP64 vector."""
    p = unpack(np.asarray(theta, np.float64))
    x = np.asarray(x, np.float64).copy()
    x[:, 7:] = 0.0
    out, zs = forward(theta, x)
    g = {}
    d = 2.0 * (out - y) / out.size            # d loss / d out
    acts = [x] + [_silu(z) for z in zs]       # inputs of layers 1..5
    g["W5"] = d.T @ acts[4]
    g["b5"] = d.sum(0)
    delta = (d @ p["W5"]) * _silu_grad(zs[3])
    for k in range(4, 0, -1):
        g[f"W{k}"] = delta.T @ acts[k - 1]
        g[f"b{k}"] = delta.sum(0)
        if k > 1:
            delta = (delta @ p[f"W{k}"]) * _silu_grad(zs[k - 2])
    g["W1"][:, 7:] = 0.0
    return np.concatenate([g[name].reshape(-1) for name, _ in SHAPES])


def round_bf16(x):
    """Values of x (float32) rounded to the nearest bfloat16, ties to even, returned as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def infer_bf16(theta, rows, in_shift, in_scale, out_shift, out_scale):
    """The matrix-core inference restated: raw feature rows (n, >= 7) -> outputs (n, 4)."""
    p = unpack(np.asarray(theta, np.float32))
    rows = np.asarray(rows, np.float32)
    x = np.zeros((len(rows), 16), np.float32)
    x[:, :7] = (rows[:, :7] - np.float32(in_shift[:7])) * np.float32(in_scale[:7])
    a = round_bf16(x)
    for k in range(1, 5):
        W = round_bf16(p[f"W{k}"]).astype(np.float64)
        if k == 1:
            W = np.pad(W, ((0, 0), (0, 8)))
        z = (a.astype(np.float64) @ W.T + p[f"b{k}"]).astype(np.float32)
        a = round_bf16(z / (np.float32(1.0) + np.exp(-z)))
    W5 = round_bf16(p["W5"]).astype(np.float64)
    y = (a.astype(np.float64) @ W5.T + p["b5"]).astype(np.float32)
    return y * np.float32(out_scale) + np.float32(out_shift)
