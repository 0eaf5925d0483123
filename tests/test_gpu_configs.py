"""The benchmark configurations' graphs on the GPU, at the sizes their checks allow.

The R-MAT configurations (vertex labels Zipf(s) over L labels, uniform edge labels) keep
their generator, label and edge-label distributions; the graph is scaled down and the
support threshold scaled with it until the check finishes:

* against the oracle at each configuration's own k (5 or 6) on R-MAT scale 7-9, and at
  k = 3 on scale 10 with low thresholds (many patterns per level);
* at R-MAT scale 16 and k = 2, without the oracle: batched equals plain, every support
  obeys MNI's anti-monotonicity and label bounds, and `extended` equals a closed form --
  the children of a single-edge row u->v (labels a <= b, edge label e) are the neighbours
  w of v with label >= a other than u (gspan.cu:447-462), plus the neighbours w of u with
  label >= a other than v whose (edge label, label) passes e < el || (e == el && b <= l)
  (gspan.cu:464-483); no backward child exists below two edges.
* a level with more than 2^21 live child sets (the set counts run in one launch of fewer
  than 2^32 work-items), batched against plain and the oracle.
"""
import numpy as np
import pytest

from gspan_amd.graph import synthetic
from gspan_amd.miner import Miner

pytestmark = pytest.mark.gpu

SPECS = {  # labels, Zipf s, edge labels, edge factor, k  (BASELINE configurations 2-5)
    "cfg2": (20, 1.1, 2, 8, 5),
    "cfg3": (100, 1.0, 4, 8, 5),
    "cfg4": (50, 1.2, 4, 16, 6),
    "cfg5": (200, 1.2, 8, 16, 6),
}


def _graph(cfg, scale):
    nl, z, el, ef, _ = SPECS[cfg]
    return synthetic("rmat", scale=scale, edge_factor=ef, n_vlabels=nl, vlabel_zipf=z,
                     n_elabels=el, seed=1)


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _predicted_weights(g, k, ms):
    with Miner(g, mode="batched", two_pass=True) as ml:
        ml.predictor_log(True, 1 << 18)
        ml.mine(k, ms)
        if ml.predictor_samples() == 0:
            return None
        ml.predictor_train(epochs=100, lr=3e-3)
        return ml.predictor_weights()


def _all_modes_match(oracle_mod, g, k, ms, want, st):
    for mode, groups in (("batched", "on"), ("batched", "off"), ("plain", "auto"),
                         ("predicted", "auto")):
        with Miner(g, mode=mode, groups=groups) as mn:
            if mode == "predicted":
                w = _predicted_weights(g, k, ms)
                if w is not None:
                    mn.set_predictor_weights(w)
            assert mn.mine(k, ms) == want, mode
            s = mn.stats()
            assert s["extended"] == st["extended"], mode
            assert s["candidates"] == st["candidates"], mode


# scale, min_sup: the oracle finishes in seconds at the configuration's own k, and frequent
# patterns reach k edges
OWN_K = {"cfg2": (8, 16), "cfg3": (9, 12), "cfg4": (8, 32), "cfg5": (7, 6)}


@pytest.mark.parametrize("cfg", sorted(OWN_K))
def test_downsized_config_own_k_matches_oracle(oracle_mod, cfg):
    scale, ms = OWN_K[cfg]
    k = SPECS[cfg][4]
    g = _graph(cfg, scale)
    _announce(f"{cfg} at R-MAT scale {scale}, k={k}, min_sup {ms}")
    want, st = oracle_mod.fsm(g, k, ms)
    assert max(len(c) for c, _ in want) == k
    _all_modes_match(oracle_mod, g, k, ms, want, st)


@pytest.mark.parametrize("cfg", ["cfg2", "cfg3", "cfg4"])
def test_downsized_config_k3_matches_oracle(oracle_mod, cfg):
    g = _graph(cfg, 10)
    ms = {"cfg2": 6, "cfg3": 3, "cfg4": 12}[cfg]
    _announce(f"{cfg} at R-MAT scale 10, k=3, min_sup {ms}")
    want, st = oracle_mod.fsm(g, 3, ms)
    assert len(want) > 1000
    _all_modes_match(oracle_mod, g, 3, ms, want, st)


def _closed_form_extended_k2(g, ms):
    """Child rows of the frequent single-edge codes, and their MNI supports, in closed form."""
    n = g.nr_nodes
    e = g.edges.astype(np.int64)
    vl = g.vlabels.astype(np.int64)
    VL, EL = int(vl.max()) + 1, int(e[:, 2].max()) + 1
    u, v, el = e[:, 0], e[:, 1], e[:, 2]
    lu, lv = vl[u], vl[v]
    # H[x, el, l]: neighbours of x with edge label el and label l; S: suffix sums over l
    H = np.zeros((n, EL, VL + 1), np.int64)
    np.add.at(H, (u, el, lv), 1)
    S = np.flip(np.cumsum(np.flip(H, 2), 2), 2)  # S[x, el, l] = sum over l' >= l
    Sall = S.sum(1)  # any edge label
    seed = lu <= lv
    u, v, el, a, b = u[seed], v[seed], el[seed], lu[seed], lv[seed]
    first = Sall[v, a] - 1  # neighbours of v with label >= a, without u
    # neighbours of u past (e, b): larger edge label with label >= a, or equal edge label
    # with label >= b; without v
    later = np.zeros(u.shape[0], np.int64)
    for x in range(EL):
        later += np.where(x > el, S[u, x, a], 0)
    later += S[u, el, b] - 1
    kids = first + later
    key = (a * EL + el) * VL + b
    sup = {}
    for kk in np.unique(key):
        sel = key == kk
        sup[int(kk)] = min(np.unique(u[sel]).shape[0], np.unique(v[sel]).shape[0])
    freq = np.array([sup[int(x)] >= ms for x in key])
    return int(kids[freq].sum()), sup


@pytest.mark.parametrize("cfg,ms", [("cfg2", 125), ("cfg3", 31)])
def test_scale16_k2_closed_form(cfg, ms):
    g = _graph(cfg, 16)
    _announce(f"{cfg} at R-MAT scale 16, k=2, min_sup {ms}: closed form")
    ext, seed_sup = _closed_form_extended_k2(g, ms)
    res = {}
    for mode, te in (("batched", "auto"), ("batched", "off"), ("plain", "auto")):
        with Miner(g, mode=mode, two_edge=te) as mn:
            res[mode, te] = (mn.mine(2, ms), mn.stats())
    pats, st = res["batched", "auto"]
    for r in res.values():
        assert r[0] == pats
        assert r[1]["extended"] == ext
    nl, _, EL, _, _ = SPECS[cfg]
    label_count = np.bincount(g.vlabels, minlength=nl)
    parent = {}
    for code, sup in pats:
        assert sup >= ms
        f, t, fl, el, tl = code[0]
        if len(code) == 1:
            assert sup == seed_sup[(fl * EL + el) * nl + tl]
            parent[code] = sup
        else:
            assert sup <= parent[code[:1]]  # MNI is anti-monotone
            labels = [fl, tl] + [x[4] for x in code[1:] if x[4] >= 0]
            assert sup <= min(label_count[x] for x in labels)
    assert len(parent) > 100 and len(pats) > len(parent)


def test_many_live_sets_batched_plain_oracle():
    """R-MAT scale 12 with 100 labels at k = 3, min_sup 4: the last batched level holds about
    6.5M live child sets, so their counts take many passes of the grid-stride count kernel.
    The oracle's report for this run (95 s of CPU) is pinned by its SHA-256 digest and
    counters in tests/golden/report_digests.json (tests/golden/make_digests.py)."""
    import hashlib
    import json
    import os
    from gspan_amd.miner import format_report
    gold = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                       "report_digests.json")))["rmat12_cfg3_k3_ms4"]
    g = synthetic(**gold["graph"])
    k, ms = gold["k"], gold["min_sup"]
    _announce("rmat12 100 labels k=3 min_sup 4: > 2^21 live sets")
    with Miner(g, mode="batched") as mn:
        got = mn.mine(k, ms)
        st = mn.stats()
    assert st["canonical"] > (1 << 21)
    assert (len(got), st["extended"], st["canonical"], st["candidates"]) == \
        (gold["patterns"], gold["extended"], gold["canonical"], gold["candidates"])
    assert hashlib.sha256(format_report(got).encode()).hexdigest() == gold["sha256"]
    with Miner(g, mode="plain") as mp:
        assert mp.mine(k, ms) == got
        assert mp.stats()["extended"] == st["extended"]
