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
