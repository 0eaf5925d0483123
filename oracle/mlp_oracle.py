"""CPU oracle of the child-count predictor's network (numpy, float64).

TEST INFRASTRUCTURE ONLY.  The predictor has no counterpart in the reference (it replaces
the size-finding count pass, gspan.cu:737-755, by a prediction with an exact fallback), so
this restates the network the library trains: MLP 32 -> 128 -> 128 -> 1 with ReLU, weights
rounded to bf16 (round to nearest even) and biases kept in fp32, mean squared error loss.
Pinned by central finite differences (tests/test_mlp_oracle.py).
"""
from __future__ import annotations

import numpy as np

F, H = 32, 128
OFF = {}
_o = 0
for name, size in (("W1", F * H), ("b1", H), ("W2", H * H), ("b2", H), ("W3", H), ("b3", 1)):
    OFF[name] = (_o, _o + size)
    _o += size
N_PARAMS = _o  # 20865


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float32 values to bf16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32)


def unpack(w: np.ndarray, round_weights: bool = True):
    w = np.asarray(w, dtype=np.float32)
    g = {k: w[a:b] for k, (a, b) in OFF.items()}
    r = bf16 if round_weights else (lambda x: x)
    return (r(g["W1"]).reshape(F, H).astype(np.float64), g["b1"].astype(np.float64),
            r(g["W2"]).reshape(H, H).astype(np.float64), g["b2"].astype(np.float64),
            r(g["W3"]).reshape(H, 1).astype(np.float64), g["b3"].astype(np.float64))


def forward(w, x, round_weights: bool = True):
    W1, b1, W2, b2, W3, b3 = unpack(w, round_weights)
    x = np.asarray(x, dtype=np.float64)
    h1 = np.maximum(x @ W1 + b1, 0.0)
    h2 = np.maximum(h1 @ W2 + b2, 0.0)
    y = (h2 @ W3)[:, 0] + b3[0]
    return y, (x, h1, h2)


def loss_and_grads(w, x, t, round_weights: bool = True):
    """Mean squared error and its gradient w.r.t. every parameter (gradients flow through the
    rounded weights unchanged, as in the library's training step)."""
    W1, b1, W2, b2, W3, b3 = unpack(w, round_weights)
    y, (x, h1, h2) = forward(w, x, round_weights)
    t = np.asarray(t, dtype=np.float64)
    n = y.shape[0]
    d = y - t
    loss = float(np.sum(d * d) / n)
    dy = (2.0 * d / n)[:, None]
    gW3 = h2.T @ dy
    gb3 = dy.sum(axis=0)
    dh2 = (dy @ W3.T) * (h2 > 0)
    gW2 = h1.T @ dh2
    gb2 = dh2.sum(axis=0)
    dh1 = (dh2 @ W2.T) * (h1 > 0)
    gW1 = x.T @ dh1
    gb1 = dh1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g
