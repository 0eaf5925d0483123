"""The learned child-count predictor on the GPU.

* the network's forward pass and gradients (the library's own kernels) match the numpy
  oracle (oracle/mlp_oracle.py) at bf16 weights with fp32 accumulation;
* training on counts logged by exact count passes lowers the loss;
* mining with predicted sizes reports exactly the oracle's patterns and supports, with
  untrained or deliberately useless weights too (every misprediction is redone exactly);
* what the trained network gains: predicted steps skip the count pass, so fewer work items
  are scanned than in the two-pass method, with identical results.
"""
import numpy as np
import pytest

import mlp_oracle as M
from gspan_amd.graph import synthetic
from gspan_amd.miner import Miner

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


@pytest.fixture(scope="module")
def small_graph():
    return synthetic("er", n=3000, m=12000, n_vlabels=4, n_elabels=2, seed=4)


def test_forward_matches_oracle(small_graph):
    _announce("mlp forward vs numpy")
    rng = np.random.default_rng(0)
    w = rng.normal(0, 0.15, M.N_PARAMS).astype(np.float32)
    x = rng.normal(0, 1, (300, M.F)).astype(np.float32)
    with Miner(small_graph) as mn:
        mn.set_predictor_weights(w)
        y = mn.mlp_forward(x)
    want, _ = M.forward(w, x)
    # fp32 accumulation over 32 + 128 + 128 terms against float64
    assert np.allclose(y, want, rtol=1e-4, atol=1e-4 * np.abs(want).max())


def test_gradients_match_oracle(small_graph):
    _announce("mlp gradients vs numpy")
    rng = np.random.default_rng(1)
    w = rng.normal(0, 0.15, M.N_PARAMS).astype(np.float32)
    x = rng.normal(0, 1, (500, M.F)).astype(np.float32)
    t = rng.normal(0, 1, 500).astype(np.float32)
    with Miner(small_graph) as mn:
        mn.set_predictor_weights(w)
        g, loss = mn.mlp_gradients(x, t)
    want_loss, want_g = M.loss_and_grads(w, x, t)
    assert np.isclose(loss, want_loss, rtol=1e-4)
    scale = np.abs(want_g).max()
    assert np.allclose(g, want_g, rtol=1e-3, atol=1e-4 * scale)


def test_training_and_predicted_mining(small_graph, oracle_mod):
    g = small_graph
    want, st = oracle_mod.fsm(g, 4, 20)
    _announce("predicted mining: untrained / useless weights fall back exactly")
    with Miner(g, mode="predicted") as mn:
        mn.set_predictor_weights(np.zeros(M.N_PARAMS, np.float32))  # predicts 0 everywhere
        assert mn.mine(4, 20) == want
        s = mn.stats()
        assert s["fallbacks"] > 0 and s["extended"] == st["extended"]
    _announce("log exact counts, train, mine with predicted sizes")
    with Miner(g, mode="batched", two_pass=True) as mb:
        mb.predictor_log(True, 1 << 18)
        assert mb.mine(4, 20) == want
        n = mb.predictor_samples()
        assert n > 1000
        w0 = mb.predictor_weights()
        l_first = mb.predictor_train(epochs=1, lr=3e-3)
        l_last = mb.predictor_train(epochs=300, lr=3e-3)
        assert l_last < 0.5 * l_first
        w = mb.predictor_weights()
        assert not np.array_equal(w, w0)
        base = mb.stats()
    with Miner(g, mode="predicted", two_pass=True) as mp:
        mp.set_predictor_weights(w)
        mp.reset_stats()
        assert mp.mine(4, 20) == want
        s = mp.stats()
    assert s["extended"] == st["extended"]
    assert s["predicted_steps"] > 0
    # the trained sizes let most steps skip the count pass
    assert s["count_launches"] < base["count_launches"]


@pytest.mark.parametrize("b3", [1e4, float("inf"), float("nan"), 30.0, 3.0])
def test_extreme_predictions_stay_exact(small_graph, oracle_mod, b3):
    """Weights whose network outputs the constant b3 for every key (all other parameters
    zero).  Huge, infinite and NaN outputs clamp to 2^30 rows per key, which no step can
    hold, so every step takes the exact two-pass path; a small output (b3 = 3: about 25
    rows per key) under-sizes the large children, whose steps are redone exactly."""
    g = small_graph
    want, st = oracle_mod.fsm(g, 4, 20)
    w = np.zeros(M.N_PARAMS, np.float32)
    w[-1] = b3
    _announce(f"predicted mining with constant output {b3}")
    with Miner(g, mode="predicted") as mn:
        mn.set_predictor_weights(w)
        assert mn.mine(4, 20) == want
        s = mn.stats()
    assert s["extended"] == st["extended"]
    if b3 == 3.0:
        assert s["fallbacks"] > 0
    else:
        assert s["predicted_steps"] == 0 and s["fallbacks"] == 0
