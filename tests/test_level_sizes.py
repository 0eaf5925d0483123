"""Level sizes of the extension tree (oracle_level_counts / oracle_level_estimate).

The exact counts come from a full mining run at min_support 1; they are pinned against the
run's own totals (single-edge rows = CSR entries with label(from) <= label(to), gspan.cu:421;
emitted rows sum to `extended`).  The Knuth random-probe estimate -- used for the
configurations too large to mine on the CPU -- must agree with the exact counts within four
standard errors at every level.
"""
import numpy as np
import pytest

from gspan_amd.graph import synthetic

CASES = [
    ("er", dict(kind="er", n=300, m=900, n_vlabels=3, n_elabels=2, seed=3), 4),
    ("er1label", dict(kind="er", n=200, m=500, n_vlabels=1, n_elabels=1, seed=8), 4),
    ("rmat8", dict(kind="rmat", scale=8, edge_factor=8, n_vlabels=4, vlabel_zipf=1.1,
                   n_elabels=2, seed=2), 3),
]


@pytest.mark.parametrize("name,spec,k", CASES, ids=[c[0] for c in CASES])
def test_estimate_matches_exact_level_sizes(oracle_mod, name, spec, k):
    kind = spec.pop("kind") if "kind" in spec else None
    g = synthetic(kind, **spec)
    spec["kind"] = kind
    rows, emitted = oracle_mod.level_counts(g, k)
    vl = g.vlabels
    e = g.edges
    assert rows[1] == int((vl[e[:, 0]] <= vl[e[:, 1]]).sum())
    _, st = oracle_mod.fsm(g, k, 1)
    assert sum(emitted[2:]) == st["extended"]
    est = oracle_mod.level_estimate(g, k, 100000, seed=7)
    assert est["rows"][1] == rows[1] and est["rows_se"][1] == 0
    for m in range(2, k + 1):
        for key, exact in (("rows", rows[m]), ("emitted", emitted[m])):
            got, se = est[key][m], est[key + "_se"][m]
            assert abs(got - exact) <= 4 * se + 1e-9 * exact, (name, key, m, got, exact, se)


@pytest.mark.parametrize("cfg,scale,ms", [("cfg2", 10, 8), ("cfg3", 10, 2), ("cfg4", 10, 10),
                                          ("cfg5", 10, 3)])
def test_k2_closed_form_matches_oracle(oracle_mod, cfg, scale, ms):
    """The closed form of `extended` at k = 2 and of the single-edge supports that the
    R-MAT scale-16 GPU test checks against (tests/test_gpu_configs.py) equals the oracle's."""
    from test_gpu_configs import SPECS, _closed_form_extended_k2, _graph
    g = _graph(cfg, scale)
    ext, sup = _closed_form_extended_k2(g, ms)
    pats, st = oracle_mod.fsm(g, 2, ms)
    assert ext == st["extended"]
    nl, _, EL, _, _ = SPECS[cfg]
    seeds = [(c, s) for c, s in pats if len(c) == 1]
    assert len(seeds) > 10
    for c, s in seeds:
        assert s == sup[(c[0][2] * EL + c[0][3]) * nl + c[0][4]]
