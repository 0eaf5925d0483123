"""Pins the CPU oracle against brute force (no reference code involved).

* full mining  == every connected pattern with <= k edges whose MNI (over all of its
                  monomorphisms) reaches min_sup, keyed by its exhaustive minimum DFS code
* canonical    == exhaustive minimum-DFS-code enumeration (codes whose first edge the
                  seeding rule gspan.cu:421 can produce)
* extension    == the child lists of a canonical code are exactly the monomorphisms of
                  each child code
* MNI          == closed forms on cliques and stars, both MNI semantics
"""
import numpy as np
import pytest

import brute
from gspan_amd.graph import csr_from_undirected


def _mine_check(oracle_mod, g, k, ms, lit):
    pats, st = oracle_mod.fsm(g, k, ms, mni_mode=1 if lit else 0, nthreads=2)
    got = {c: s for c, s in pats}
    want = brute.frequent_patterns(g, k, ms, omp_literal=lit)
    assert got == want
    assert st["overruns"] == 0
    assert len(got) == len(pats)  # no pattern reported twice


@pytest.mark.parametrize("seed", range(12))
def test_mining_matches_brute_force(oracle_mod, seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(5, 10))
    m = int(rng.integers(n - 1, min(15, n * (n - 1) // 2)))
    g = brute.random_small_graph(rng, n, m, int(rng.integers(1, 4)), int(rng.integers(1, 3)))
    k = 4 if m <= 11 else 3
    ms = int(rng.integers(1, 4))
    _mine_check(oracle_mod, g, k, ms, lit=False)
    _mine_check(oracle_mod, g, k, ms, lit=True)


def test_canonical_matches_exhaustive_minimum(oracle_mod):
    rng = np.random.default_rng(7)
    checked = canon = 0
    for _ in range(3000):
        code = brute.random_code(rng, 5, int(rng.integers(1, 4)), int(rng.integers(1, 3)))
        if len(code) == 1 or code[0][2] > code[0][4]:
            continue  # the seeding rule never produces a first edge with fl > tl
        want = brute.is_min_code(code)
        assert oracle_mod.is_canonical(code) == int(want), code
        checked += 1
        canon += want
    assert checked > 1500 and 100 < canon < checked


def test_extension_rows_are_all_monomorphisms(oracle_mod):
    rng = np.random.default_rng(3)
    for _ in range(6):
        g = brute.random_small_graph(rng, 8, 13, 2, 2)
        sg = brute.SmallGraph(g)
        # walk two levels down from every seed
        for e5, rows in oracle_mod.seed(g):
            code = (e5,)
            layers = [rows]
            kids = oracle_mod.extend(g, code, layers, nthreads=2)
            for ce5, crows in kids:
                child = code + (ce5,)
                # rebuild each child row's entry sequence and compare with brute force
                seqs = sorted((int(rows[p][1]), int(eid)) for p, eid in crows)
                want = sorted(brute.monomorphisms(sg, child))
                assert seqs == want, child
                assert len(set(seqs)) == len(seqs)


def _clique(n, label=0):
    u, v = np.triu_indices(n, 1)
    return csr_from_undirected(n, u, v, np.zeros(u.shape[0], np.int64), np.full(n, label))


def test_mni_closed_form_clique(oracle_mod):
    # every connected pattern of K_n with one label has MNI = n
    g = _clique(6)
    pats, _ = oracle_mod.fsm(g, 4, 1)
    assert pats and all(s == 6 for _, s in pats)
    pats_l, _ = oracle_mod.fsm(g, 4, 1, mni_mode=1)
    assert all(s == 6 for _, s in pats_l)


def test_mni_closed_form_star(oracle_mod):
    # star K_{1,s}, centre label 0 < leaf label 1: the single edge (0,1,0,0,1) has
    # GPU-semantics MNI min(1, s) = 1, while the literal OpenMP semantics never count
    # vertex 0 and give s
    s = 7
    g = csr_from_undirected(s + 1, np.zeros(s, np.int64), np.arange(1, s + 1),
                            np.zeros(s, np.int64), np.array([0] + [1] * s))
    pats, _ = oracle_mod.fsm(g, 1, 1)
    assert pats == [(((0, 1, 0, 0, 1),), 1)]
    pats_l, _ = oracle_mod.fsm(g, 1, 1, mni_mode=1)
    assert pats_l == [(((0, 1, 0, 0, 1),), s)]
    # two-edge star pattern: GPU semantics 1 (centre), literal counts the centre too
    # (edge 1 leaves vertex 0) -> 1
    pats2, _ = oracle_mod.fsm(g, 2, 1)
    assert (((0, 1, 0, 0, 1), (0, 2, -1, 0, 1)), 1) in pats2


def test_mni_antimonotone_and_label_bound(oracle_mod):
    from gspan_amd.graph import synthetic
    g = synthetic("er", n=300, m=900, n_vlabels=3, n_elabels=2, seed=5)
    pats, _ = oracle_mod.fsm(g, 3, 5)
    sup = {c: s for c, s in pats}
    freq = np.bincount(g.vlabels)
    for c, s in pats:
        if len(c) > 1:
            assert c[:-1] in sup and sup[c[:-1]] >= s
        labels = {c[0][2], c[0][4]} | {e[4] for e in c if e[4] >= 0}
        assert s <= min(freq[l] for l in labels)


def test_report_order_is_preorder(oracle_mod):
    # every pattern's parent precedes it; siblings follow the reference's child order
    from gspan_amd.graph import synthetic
    g = synthetic("er", n=200, m=700, n_vlabels=2, n_elabels=2, seed=9)
    pats, _ = oracle_mod.fsm(g, 3, 3)
    pos = {c: i for i, (c, _) in enumerate(pats)}
    for c, _ in pats:
        if len(c) > 1:
            assert pos[c[:-1]] < pos[c]
    def key(code):
        out = [code[0][2:5]]
        for a, b, _fl, el, tl in code[1:]:
            out.append((0, b, el, 0) if a > b else (1, -a, el, tl))
        return out
    assert [c for c, _ in pats] == sorted((c for c, _ in pats), key=key)
