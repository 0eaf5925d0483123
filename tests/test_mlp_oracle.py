"""Pins the CPU oracle of the predictor's network: analytic gradients against central
finite differences (float64, unrounded weights), and bf16 rounding against known values."""
import numpy as np

import mlp_oracle as M


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.140625, 0.0, 65504.0], np.float32)
    # 1 + 2^-8 is halfway between bf16 1.0 and 1.0078125: ties go to the even mantissa (1.0)
    want = np.array([1.0, 1.0, 1.0078125, -3.140625, 0.0, 65536.0], np.float32)
    assert np.array_equal(M.bf16(x), want)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    w = rng.normal(0, 0.2, M.N_PARAMS).astype(np.float64)
    x = rng.normal(0, 1, (7, M.F))
    t = rng.normal(0, 1, 7)
    loss, g = M.loss_and_grads(w, x, t, round_weights=False)
    idx = rng.choice(M.N_PARAMS, 60, replace=False)
    eps = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += eps
        wm[i] -= eps
        # float32 unpacking would hide the step: evaluate in float64 directly
        lp = _loss64(wp, x, t)
        lm = _loss64(wm, x, t)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-5 + 1e-4 * abs(fd), (i, fd, g[i])


def _loss64(w, x, t):
    o = M.OFF
    W1 = w[o["W1"][0]:o["W1"][1]].reshape(M.F, M.H)
    b1 = w[o["b1"][0]:o["b1"][1]]
    W2 = w[o["W2"][0]:o["W2"][1]].reshape(M.H, M.H)
    b2 = w[o["b2"][0]:o["b2"][1]]
    W3 = w[o["W3"][0]:o["W3"][1]].reshape(M.H, 1)
    b3 = w[o["b3"][0]]
    h1 = np.maximum(x @ W1 + b1, 0)
    h2 = np.maximum(h1 @ W2 + b2, 0)
    y = (h2 @ W3)[:, 0] + b3
    return float(np.mean((y - t) ** 2))


def test_loss_matches_definition():
    rng = np.random.default_rng(1)
    w = rng.normal(0, 0.1, M.N_PARAMS).astype(np.float32)
    x = rng.normal(0, 1, (5, M.F))
    t = rng.normal(0, 1, 5)
    y, _ = M.forward(w, x)
    loss, _ = M.loss_and_grads(w, x, t)
    assert np.isclose(loss, np.mean((y - t) ** 2))
