"""GPU parity: every kernel path against the CPU oracle on the same seeded inputs.

Integer graph work, so the bar is bit-exact: identical pattern sets, identical MNI
supports, identical child row multisets, identical canonicality flags.
"""
import numpy as np
import pytest

import brute
from gspan_amd.graph import CSRGraph, synthetic, csr_from_undirected
from gspan_amd import _lib as L
from gspan_amd.miner import Miner, format_report

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def small_graphs(count, seed0=0):
    for s in range(count):
        rng = np.random.default_rng(1000 + seed0 + s)
        n = int(rng.integers(6, 14))
        m = int(rng.integers(n - 1, min(24, n * (n - 1) // 2)))
        yield s, rng, brute.random_small_graph(rng, n, m, int(rng.integers(1, 4)),
                                               int(rng.integers(1, 3)))


def test_canonical_batch_matches_oracle(oracle_mod):
    _announce("canonical batch vs oracle")
    g = synthetic("er", n=50, m=100, n_vlabels=2, seed=1)
    rng = np.random.default_rng(5)
    codes = []
    while len(codes) < 2000:
        c = brute.random_code(rng, 6, int(rng.integers(1, 4)), int(rng.integers(1, 3)))
        if c[0][2] <= c[0][4]:
            codes.append(c)
    with Miner(g) as mn:
        got = mn.is_canonical(codes)
    want = [oracle_mod.is_canonical(c) == 1 for c in codes]
    assert got == want
    assert 50 < sum(want) < len(want)


def test_canonical_two_edge_codes_exhaustive(oracle_mod):
    """Every two-edge code over three vertex labels and two edge labels (first edges in
    both label orders; forward second edges from either vertex, backward second edges)
    against the oracle: the kernels walk two-edge codes in registers."""
    _announce("canonical two-edge codes, exhaustive")
    g = synthetic("er", n=60, m=150, n_vlabels=3, n_elabels=2, seed=2)
    codes = []
    for a in range(3):
        for e0 in range(2):
            for b in range(3):
                first = (0, 1, a, e0, b)
                for e in range(2):
                    codes.append([first, (1, 0, -1, e, -1)])
                    for s in range(2):
                        for l in range(3):
                            codes.append([first, (s, 2, -1, e, l)])
    with Miner(g) as mn:
        got = mn.is_canonical(codes)
    want = [oracle_mod.is_canonical(c) == 1 for c in codes]
    assert got == want
    assert 20 < sum(want) < len(want)


def test_canonical_three_edge_codes(oracle_mod):
    """Random three-edge codes (paths, stars, triangles, in every DFS order, canonical or
    not) against the oracle: the kernels walk codes of three edges with one thread."""
    _announce("canonical three-edge codes")
    g = synthetic("er", n=60, m=150, n_vlabels=3, n_elabels=2, seed=3)
    rng = np.random.default_rng(11)
    codes = []
    while len(codes) < 3000:
        c = brute.random_code(rng, 3, int(rng.integers(1, 4)), int(rng.integers(1, 3)))
        if len(c) == 3:
            codes.append(c)
    with Miner(g) as mn:
        got = mn.is_canonical(codes)
    want = [oracle_mod.is_canonical(c) == 1 for c in codes]
    assert got == want
    assert 100 < sum(want) < len(want)


def test_canonical_star_projection_count(oracle_mod):
    # K_{1,6} with one label: its minimum code starts at a leaf and walks the other five
    # leaves from the centre; the walk carries up to 6! = 720 projections (cap 2048).
    # Rooting the code at the centre is not canonical.
    _announce("canonical 6-star")
    g = synthetic("er", n=50, m=100, n_vlabels=1, seed=1)
    leaf = ((0, 1, 0, 0, 0),) + tuple((1, j, -1, 0, 0) for j in range(2, 7))
    centre = ((0, 1, 0, 0, 0),) + tuple((0, j, -1, 0, 0) for j in range(2, 7))
    with Miner(g) as mn:
        assert mn.is_canonical([leaf, centre]) == [True, False]
    assert oracle_mod.is_canonical(leaf) == 1 and oracle_mod.is_canonical(centre) == 0


def test_canonical_seven_star(oracle_mod):
    """K_{1,7} with one label: the walk of its minimum code (rooted at a leaf) carries
    7! = 5040 projections, the most any code of up to 7 edges has.  Mining the star graph
    K_{1,7} itself to k = 7 checks that 7-edge candidate in both modes."""
    _announce("canonical 7-star")
    g = synthetic("er", n=50, m=100, n_vlabels=1, seed=1)
    leaf = ((0, 1, 0, 0, 0),) + tuple((1, j, -1, 0, 0) for j in range(2, 8))
    centre = ((0, 1, 0, 0, 0),) + tuple((0, j, -1, 0, 0) for j in range(2, 8))
    with Miner(g) as mn:
        assert mn.is_canonical([leaf, centre, leaf[:6]]) == [True, False, True]
    assert oracle_mod.is_canonical(leaf) == 1 and oracle_mod.is_canonical(centre) == 0
    star = _star(7)
    want, st = oracle_mod.fsm(star, 7, 1)
    assert want[-1] == (leaf, 1)
    for mode in ("batched", "plain"):
        _announce(f"mine K_1,7 k=7 mode={mode}")
        with Miner(star, mode=mode) as mn:
            assert mn.mine(7, 1) == want, mode
            assert mn.stats()["extended"] == st["extended"], mode


@pytest.mark.parametrize("mni", ["gpu", "omp-literal"])
def test_support_matches_oracle(oracle_mod, mni):
    for s, rng, g in small_graphs(6):
        _announce(f"support mni={mni} graph={s}")
        with Miner(g, mni=mni) as mn:
            for e5, rows in oracle_mod.seed(g):
                code, layers = (e5,), [rows]
                mode = 1 if mni == "omp-literal" else 0
                assert mn.support(code, layers) == oracle_mod.support(g, code, layers, mode)
                for ce5, crows in oracle_mod.extend(g, code, layers, nthreads=1)[:4]:
                    c2, l2 = code + (ce5,), layers + [crows]
                    assert mn.support(c2, l2) == oracle_mod.support(g, c2, l2, mode)


def _canon_rows(rows):
    return sorted(map(tuple, np.asarray(rows).tolist()))


def test_extend_matches_oracle(oracle_mod):
    for s, rng, g in small_graphs(6, seed0=50):
        _announce(f"extend graph={s}")
        with Miner(g) as mn:
            for e5, rows in oracle_mod.seed(g)[:6]:
                frontier = [((e5,), [rows])]
                for depth in range(3):
                    nxt = []
                    for code, layers in frontier[:4]:
                        want = oracle_mod.extend(g, code, layers, nthreads=2)
                        got = mn.extend(code, layers)
                        assert [w[0] for w in want] == [x[0] for x in got], code
                        for (we5, wrows), (ge5, grows, gcanon, gsup) in zip(want, got):
                            assert _canon_rows(wrows) == _canon_rows(grows), (code, we5)
                            child = code + (we5,)
                            assert gcanon == (oracle_mod.is_canonical(child) == 1)
                            assert gsup == oracle_mod.support(g, child, layers + [wrows], 0)
                            nxt.append((child, layers + [wrows]))
                    frontier = nxt


@pytest.mark.parametrize("mode", ["batched", "plain"])
@pytest.mark.parametrize("mni", ["gpu", "omp-literal"])
def test_mining_small_graphs(oracle_mod, mode, mni):
    for s, rng, g in small_graphs(8, seed0=200):
        k = int(rng.integers(2, 5))
        ms = int(rng.integers(1, 4))
        _announce(f"mine small mode={mode} mni={mni} graph={s} k={k} ms={ms}")
        want, st = oracle_mod.fsm(g, k, ms, mni_mode=1 if mni == "omp-literal" else 0)
        with Miner(g, mode=mode, mni=mni) as mn:
            got = mn.mine(k, ms)
            gst = mn.stats()
        assert got == want
        assert gst["extended"] == st["extended"]
        assert gst["candidates"] == st["candidates"]
        assert gst["canonical"] == st["canonical"]


def test_mining_er_config_k3(oracle_mod):
    g = synthetic("er", n=10000, m=40000, n_vlabels=5, n_elabels=1, seed=1)
    _announce("mine ER 10k/40k k=3 min_sup=200")
    want, st = oracle_mod.fsm(g, 3, 200)
    with Miner(g) as mn:
        got = mn.mine(3, 200)
        assert mn.stats()["extended"] == st["extended"]
    assert format_report(got) == oracle_mod.format_report(want)


def test_mining_er_bench_config_all_modes(oracle_mod):
    """The bench workload itself (ER 10,000 / 40,000, 5 labels, min_sup 200, k = 4): the
    report text and the extended-embedding total equal the oracle's in every mode."""
    g = synthetic("er", n=10000, m=40000, n_vlabels=5, n_elabels=1, seed=1)
    _announce("mine ER 10k/40k k=4 min_sup=200 (bench config), batched/plain/two-pass")
    want, st = oracle_mod.fsm(g, 4, 200)
    text = oracle_mod.format_report(want)
    assert len(want) == 4415
    for kw in (dict(mode="batched"), dict(mode="plain"), dict(mode="batched", two_pass=True),
               dict(mode="batched", sets="bits")):
        with Miner(g, **kw) as mn:
            got = mn.mine(4, 200)
            assert mn.stats()["extended"] == st["extended"], kw
        assert format_report(got) == text, kw


def test_mining_rmat_small(oracle_mod):
    g = synthetic("rmat", scale=11, edge_factor=8, n_vlabels=8, vlabel_zipf=1.1, n_elabels=2,
                  seed=1)
    _announce("mine rmat scale 11 k=3 min_sup=30")
    want, st = oracle_mod.fsm(g, 3, 30)
    for mode in ("batched", "plain"):
        with Miner(g, mode=mode) as mn:
            got = mn.mine(3, 30)
            assert mn.stats()["extended"] == st["extended"]
        assert got == want, mode


def test_chunked_and_split_runs_agree(oracle_mod):
    g = synthetic("er", n=3000, m=12000, n_vlabels=4, n_elabels=2, seed=4)
    _announce("mine ER 3k chunked / budget-split / two-pass")
    want, st = oracle_mod.fsm(g, 3, 20)
    with Miner(g, chunk_rows=1000) as mn:
        assert mn.mine(3, 20) == want
    with Miner(g, budget_bytes=48 << 20) as mn:
        assert mn.mine(3, 20) == want
    for mode in ("batched", "plain"):
        with Miner(g, mode=mode, two_pass=True) as mn:
            assert mn.mine(3, 20) == want
            s = mn.stats()
            assert s["extended"] == st["extended"] and s["count_launches"] > 0


@pytest.mark.parametrize("sets", ["bits", "bytes"])
def test_set_kinds_match_oracle(oracle_mod, sets):
    """MNI image sets as bitmaps (atomic OR) and as bytemaps (plain byte stores), in the
    single-pass last level and in the two-pass path, give the oracle's supports."""
    g = synthetic("rmat", scale=11, edge_factor=8, n_vlabels=6, vlabel_zipf=1.1, n_elabels=2,
                  seed=7)
    _announce(f"mine rmat scale 11 k=3 min_sup=40 sets={sets}")
    want, st = oracle_mod.fsm(g, 3, 40)
    for mode in ("batched", "plain"):
        for two_pass in (False, True):
            with Miner(g, mode=mode, sets=sets, two_pass=two_pass) as mn:
                assert mn.mine(3, 40) == want, (mode, two_pass)
                assert mn.stats()["extended"] == st["extended"]


def test_tile_shapes_and_tuner(oracle_mod):
    """Every tile size of the extension kernels gives the oracle's report; the tuner times
    each launch shape once per bucket and then keeps the fastest, without changing results."""
    g = synthetic("er", n=10000, m=40000, n_vlabels=5, n_elabels=1, seed=1)
    _announce("ER 10k k=3 tiles 256/512/1024 + tuner")
    want, st = oracle_mod.fsm(g, 3, 200)
    for tile in (256, 512, 1024):
        with Miner(g, tile=tile) as mn:
            assert mn.mine(3, 200) == want, tile
            assert mn.stats()["extended"] == st["extended"]
    with Miner(g, autotune=True) as mn:
        for _ in range(8):  # enough runs to try every shape in every bucket
            assert mn.mine(3, 200) == want
        rep = mn.tuner_report()
    assert rep, "no launch was large enough to tune"
    for name, b in rep.items():
        shapes = b["ns_per_item"]
        assert {s[0] for s in shapes} == {256, 512, 1024}, name
        tried = [s for s in shapes if s[2] >= 0]
        assert len(tried) == len(shapes), (name, shapes)
        best = min(tried, key=lambda s: s[2])
        assert b["best"] == [best[0], best[1]], (name, b)


def test_spill_rows_to_host_memory(oracle_mod):
    """With a device budget too small for any level, every level's rows live in pinned
    host memory that the kernels reach through its device mapping; results unchanged."""
    g = synthetic("er", n=2000, m=8000, n_vlabels=3, n_elabels=2, seed=9)
    _announce("mine ER 2k k=3 ms=15 with spilled level rows")
    want, st = oracle_mod.fsm(g, 3, 15)
    for mode in ("batched", "plain"):
        with Miner(g, mode=mode, budget_bytes=1) as mn:
            assert mn.mine(3, 15) == want, mode
            s = mn.stats()
            assert s["extended"] == st["extended"]
            assert s["spilled_rows"] > 0


def test_two_edge_level_with_and_without_early_pass(oracle_mod):
    """Mining to k = 2 on a simple graph: by default the two-edge pass starts from the class
    sums before the host lists the single-edge codes; with a device budget too small for its
    tables the step over the codes runs its own pass.  Reports equal the oracle's either way."""
    g = synthetic("rmat", scale=11, edge_factor=8, n_vlabels=6, vlabel_zipf=1.1, n_elabels=2,
                  seed=5)
    for mni, name in ((0, "gpu"), (1, "omp-literal")):
        _announce(f"mine rmat scale 11 k=2 ms=30 budget 1 B mni={name}")
        want, st = oracle_mod.fsm(g, 2, 30, mni_mode=mni)
        assert want
        with Miner(g, mni=name, budget_bytes=1) as mn:
            assert mn.mine(2, 30) == want, name
            s = mn.stats()
            assert s["extended"] == st["extended"]
            assert s["two_edge_levels"] >= 1
        with Miner(g, mni=name) as mn:
            assert mn.mine(2, 30) == want, name
            assert mn.stats()["two_edge_levels"] == 1


def test_prune_graph_matches_oracle(oracle_mod):
    """gs_prune_graph (graph.cc:216-266 on the device) against the oracle's restatement,
    at every threshold that changes the kept set; then mining the pruned graph must give
    the unpruned graph's report (MNI support is bounded by each pattern label's count)."""
    cases = [("empty", CSRGraph(np.zeros(1, np.int32), np.zeros(0, np.int32),
                                np.zeros((0, 3), np.int32)))]
    for s in range(4):
        cases.append((f"er{s}", synthetic("er", n=200 + 50 * s, m=600 + 100 * s,
                                          n_vlabels=3 + s, n_elabels=1 + s % 2, seed=20 + s,
                                          vlabel_zipf=1.2 if s % 2 else None)))
    cases.append(("rmat12", synthetic("rmat", scale=12, edge_factor=8, n_vlabels=20,
                                      vlabel_zipf=1.1, n_elabels=2, seed=3)))
    for name, g in cases:
        _announce(f"prune_graph {name} n={g.nr_nodes} nnz={g.nr_edges}")
        run = oracle_mod.prune_graph(g)
        sups = sorted({0, 1, g.nr_nodes + 1} | set(np.bincount(g.vlabels).tolist()))
        with Miner(g) as mn:
            for ms in sups:
                got, want = mn.prune_graph(ms), run(ms)
                assert got.nr_nodes == want.nr_nodes and got.nr_edges == want.nr_edges, (name, ms)
                np.testing.assert_array_equal(got.offsets, want.offsets)
                np.testing.assert_array_equal(got.vlabels, want.vlabels)
                np.testing.assert_array_equal(got.edges, want.edges)
    g = cases[-2][1]
    counts = np.bincount(g.vlabels)
    ms = int(np.sort(counts[counts > 0])[1]) + 1  # drops at least the two rarest labels
    _announce(f"mine pruned vs unpruned ms={ms}")
    with Miner(g) as mn:
        pruned = mn.prune_graph(ms)
        want = mn.mine(3, ms)
    assert pruned.nr_nodes < g.nr_nodes
    with Miner(pruned) as mp:
        assert mp.mine(3, ms) == want


def test_prune_graph_matches_reference_golden():
    """gs_prune_graph against the reference's own prune_graph outputs (tests/golden)."""
    import glob
    import os
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    g = CSRGraph.load(os.path.join(gold, "prune_in.bin"))
    cases = []
    for p in glob.glob(os.path.join(gold, "prune_*.bin")):
        tag = os.path.basename(p)[len("prune_"):-len(".bin")]
        if tag.isdigit():
            cases.append((int(tag), p))
    assert len(cases) >= 4
    _announce(f"prune golden, {len(cases)} thresholds")
    with Miner(g) as mn:
        for ms, p in sorted(cases):
            want, got = CSRGraph.load(p), mn.prune_graph(ms)
            np.testing.assert_array_equal(got.offsets, want.offsets)
            np.testing.assert_array_equal(got.vlabels, want.vlabels)
            np.testing.assert_array_equal(got.edges, want.edges)


def test_mining_edge_cases(oracle_mod):
    """k = 1, thresholds above every support, no edges, isolated vertices, the empty graph,
    and a bitmap/bytemap budget of one byte: every case gives the oracle's report."""
    cases = []
    g = synthetic("er", n=400, m=900, n_vlabels=3, n_elabels=2, seed=31)
    cases += [("k1", g, 1, 5), ("ms_above_all", g, 3, 10**6), ("ms1_k2", g, 2, 1)]
    iso = csr_from_undirected(12, np.array([0, 1, 2, 3]), np.array([1, 2, 3, 0]),
                              np.array([0, 1, 0, 1]), np.array([0, 1] * 6))
    cases += [("isolated", iso, 3, 1)]
    noedge = CSRGraph(np.zeros(6, np.int32), np.array([0, 1, 2, 0, 1], np.int32),
                      np.zeros((0, 3), np.int32))
    cases += [("no_edges", noedge, 3, 1)]
    empty = CSRGraph(np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros((0, 3), np.int32))
    cases += [("empty", empty, 2, 1)]
    for name, gg, k, ms in cases:
        _announce(f"mine edge case {name} k={k} ms={ms}")
        want, st = oracle_mod.fsm(gg, k, ms)
        for mode in ("batched", "plain"):
            with Miner(gg, mode=mode) as mn:
                assert mn.mine(k, ms) == want, (name, mode)
                assert mn.stats()["extended"] == st["extended"], (name, mode)


def _star(d):
    """Hub 0 joined to leaves 1..d; one vertex label, one edge label."""
    leaves = np.arange(1, d + 1)
    return csr_from_undirected(d + 1, np.zeros(d, np.int64), leaves, np.zeros(d, np.int64),
                               np.zeros(d + 1, np.int64))


def test_child_count_past_int_max(oracle_mod):
    """More than 2^31 - 1 embeddings of one child.  In a star whose hub has degree d, the
    3-edge star has d(d-1)(d-2) embeddings: 2,191,922,600 for d = 1300, past a 32-bit count.
    With one label, mining to k = 3 reports the edge (support d + 1: every vertex can image
    either end), the 2-path and the 3-star (support 1: only the hub can be their centre).
    The oracle fixes the codes and order on a star of degree 40; the large star must give
    the same codes with those closed-form supports in every mode.  Extending the 2-path
    on the large star must store its child rows, more than 2^31 of them: that is an error."""
    _announce("star d=40 vs oracle")
    small = _star(40)
    want_small, _ = oracle_mod.fsm(small, 3, 1)
    assert [s for _, s in want_small] == [41, 1, 1]
    d = 1300
    assert d * (d - 1) * (d - 2) > 2**31 - 1
    want = [(code, s) for (code, _), s in zip(want_small, [d + 1, 1, 1])]
    big = _star(d)
    for mode in ("batched", "plain"):
        _announce(f"star d={d} k=3 mode={mode}")
        with Miner(big, mode=mode) as mn:
            assert mn.mine(3, 1) == want, mode
            assert mn.stats()["extended"] >= 2**31, mode  # a saturated count adds 2^31
    _announce(f"star d={d} extend 2-path")
    path = [(0, 1, 0, 0, 0), (1, 2, -1, 0, 0)]
    (e5, l0), = oracle_mod.seed(big)
    assert e5 == path[0] and len(l0) == 2 * d
    with Miner(big) as mn:
        ch = mn.extend(path[:1], [l0])
        rows = next(r for e5, r, _, _ in ch if e5 == path[1])
        assert len(rows) == d * (d - 1)
        with pytest.raises(L.GsError, match="2\\^31") as ei:
            mn.extend(path, [l0, rows])
        assert ei.value.code == L.GS_ERR_BUDGET


def test_large_key_space(oracle_mod):
    """A level whose child key space is far above what one workgroup compacts (128k keys):
    60 vertex labels and 4 edge labels at min_sup 1 make ~7,000 frequent edges with 480
    forward keys each (3.2M keys in the batched step), so the ordered compaction takes its
    multi-launch path; plain steps stay on the one-workgroup path.  Both modes give the
    oracle's report."""
    g = synthetic("er", n=3000, m=20000, n_vlabels=60, n_elabels=4, seed=5)
    want, st = oracle_mod.fsm(g, 2, 1)
    assert len(want) > 200_000
    for mode in ("batched", "plain"):
        _announce(f"large key space mode={mode}")
        with Miner(g, mode=mode) as mn:
            assert mn.mine(2, 1) == want, mode
            assert mn.stats()["extended"] == st["extended"], mode


@pytest.mark.parametrize("seed", range(6))
def test_multigraphs_all_modes(oracle_mod, seed):
    """Edge lists with repeated pairs (same or other edge label) and self-loops, which the
    reference's CSR keeps (graph.cc:32-36): such a graph is not simple, so a neighbour run's
    row vertices are looked up in the run.  Every method against the oracle."""
    rng = np.random.default_rng(7100 + seed)
    n = int(rng.integers(20, 80))
    m = int(rng.integers(2 * n, 5 * n))
    loops = rng.integers(0, n, size=max(1, n // 10))
    u = np.concatenate([rng.integers(0, n, size=m), loops])
    v = np.concatenate([rng.integers(0, n, size=m), loops])
    el = rng.integers(0, int(rng.integers(1, 3)), size=u.shape[0])
    vl = rng.integers(0, int(rng.integers(1, 4)), size=n)
    g = csr_from_undirected(n, u, v, el, vl)
    k = int(rng.integers(2, 4))
    ms = int(rng.integers(2, 8))
    mni = int(rng.integers(0, 2))
    _announce(f"multigraph seed={seed} n={n} m={u.shape[0]} k={k} ms={ms} mni={mni}")
    want, st = oracle_mod.fsm(g, k, ms, mni_mode=mni)
    for mode, sets, groups, two in (("batched", "auto", "on", False),
                                    ("batched", "auto", "off", False),
                                    ("plain", "auto", "on", False),
                                    ("batched", "bits", "on", True)):
        with Miner(g, mode=mode, mni=("gpu", "omp-literal")[mni], sets=sets, groups=groups,
                   two_pass=two) as mn:
            assert mn.mine(k, ms) == want, (mode, sets, groups, two)
            assert mn.stats()["extended"] == st["extended"], (mode, sets, groups, two)


@pytest.mark.parametrize("seed", range(12))
def test_random_graphs_all_modes(oracle_mod, seed):
    """Seeded ER and small R-MAT graphs with random label counts, k, threshold and MNI
    semantics: batched, plain and bitmap-set runs, with the last level by neighbour runs and
    row groups or by every neighbour, single-pass or two-pass, all give the oracle's report
    and `extended` total."""
    rng = np.random.default_rng(7000 + seed)
    if seed % 3:
        nv = int(rng.integers(30, 400))
        m = int(rng.integers(nv, 4 * nv))
        g = synthetic("er", n=nv, m=m, n_vlabels=int(rng.integers(1, 6)),
                      n_elabels=int(rng.integers(1, 4)), seed=seed)
    else:
        g = synthetic("rmat", scale=int(rng.integers(6, 10)), edge_factor=4,
                      n_vlabels=int(rng.integers(1, 6)), vlabel_zipf=1.1,
                      n_elabels=int(rng.integers(1, 3)), seed=seed)
    k = int(rng.integers(2, 5))
    ms = int(rng.integers(2, 12))
    mni = int(rng.integers(0, 2))
    _announce(f"random graph seed={seed} k={k} ms={ms} mni={mni}")
    want, st = oracle_mod.fsm(g, k, ms, mni_mode=mni)
    for mode, sets, groups, two, te in (("batched", "auto", "auto", False, "auto"),
                                        ("batched", "auto", "auto", False, "off"),
                                        ("plain", "auto", "on", False, "auto"),
                                        ("batched", "bits", "off", False, "off"),
                                        ("batched", "auto", "on", True, "auto"),
                                        ("plain", "auto", "off", True, "auto")):
        with Miner(g, mode=mode, mni=("gpu", "omp-literal")[mni], sets=sets, groups=groups,
                   two_pass=two, two_edge=te) as mn:
            assert mn.mine(k, ms) == want, (mode, sets, groups, two, te)
            assert mn.stats()["extended"] == st["extended"], (mode, sets, groups, two, te)
