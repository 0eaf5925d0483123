"""Data formats and generators: the reference's binary CSR layout and text conversion.

Where oracle/_ref holds the reference's own converter (built from its graph.cc and
convert.cc in this container), the Python converter and the C++ `convert` CLI must
write byte-identical files.  Those tests are skipped where the reference is absent.
"""
import os
import subprocess

import numpy as np
import pytest

import brute
from gspan_amd import graph as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONVERT = os.path.join(ROOT, "gspan_amd", "bin", "convert")


def _ref():
    import oracle
    return oracle if oracle.ref_available() else None


def test_binary_round_trip(tmp_path):
    g = G.synthetic("er", n=500, m=1500, n_vlabels=4, n_elabels=3, seed=2)
    p = tmp_path / "g.bin"
    g.save(p)
    h = G.CSRGraph.load(p)
    assert np.array_equal(g.offsets, h.offsets)
    assert np.array_equal(g.vlabels, h.vlabels)
    assert np.array_equal(g.edges, h.edges)
    # header + offsets + labels + 12-byte edges (graph.cc:202-214)
    assert os.path.getsize(p) == 4 * (2 + g.nr_nodes + 1 + g.nr_nodes) + 12 * g.nr_edges


@pytest.mark.parametrize("cut,msg", [(0, "nr_nodes"), (6, "nr_edges"), (12, "bad offsets"),
                                     (-13, "bad edges")])
def test_truncated_file_errors(tmp_path, cut, msg):
    g = G.synthetic("er", n=20, m=30, n_vlabels=2, seed=1)
    p = tmp_path / "g.bin"
    g.save(p)
    data = p.read_bytes()
    (tmp_path / "t.bin").write_bytes(data[:cut] if cut >= 0 else data[:cut])
    with pytest.raises(RuntimeError, match=msg):
        G.CSRGraph.load(tmp_path / "t.bin")


def test_missing_file_errors(tmp_path):
    with pytest.raises(OSError):
        G.CSRGraph.load(tmp_path / "nope.bin")


def test_csr_invariants_of_generators():
    for g in (G.synthetic("er", n=1000, m=4000, n_vlabels=5, seed=1),
              G.synthetic("rmat", scale=10, edge_factor=8, n_vlabels=20, vlabel_zipf=1.1,
                          n_elabels=2, seed=1)):
        e = g.edges
        assert g.offsets[0] == 0 and g.offsets[-1] == g.nr_edges
        own = np.repeat(np.arange(g.nr_nodes), np.diff(g.offsets))
        assert np.array_equal(e[:, 0], own)
        # adjacency sorted by (to, label), no self loops, symmetric
        key = e[:, 0].astype(np.int64) << 40 | e[:, 1].astype(np.int64) << 8 | e[:, 2]
        assert np.all(np.diff(key) > 0)
        assert np.all(e[:, 0] != e[:, 1])
        fwd = set(map(tuple, e.tolist()))
        assert all((b, a, w) in fwd for a, b, w in e.tolist()[:2000])


def test_generators_are_seeded():
    a = G.synthetic("rmat", scale=9, edge_factor=8, n_vlabels=10, vlabel_zipf=1.0, seed=3)
    b = G.synthetic("rmat", scale=9, edge_factor=8, n_vlabels=10, vlabel_zipf=1.0, seed=3)
    c = G.synthetic("rmat", scale=9, edge_factor=8, n_vlabels=10, vlabel_zipf=1.0, seed=4)
    assert np.array_equal(a.edges, b.edges) and np.array_equal(a.vlabels, b.vlabels)
    assert not np.array_equal(a.edges, c.edges)


def test_er_config_shape():
    g = G.synthetic("er", n=10000, m=40000, n_vlabels=5, n_elabels=1, seed=1)
    assert g.nr_nodes == 10000 and g.nr_edges == 80000
    assert set(np.unique(g.vlabels)) == set(range(5))
    assert set(np.unique(g.edges[:, 2])) == {0}


def test_zipf_label_frequencies():
    rng = np.random.default_rng(0)
    lab = G.zipf_labels(rng, 200000, 20, 1.1)
    p0 = (1 / np.arange(1, 21) ** 1.1)
    p0 = p0 / p0.sum()
    freq = np.bincount(lab, minlength=20) / lab.shape[0]
    assert np.allclose(freq, p0, atol=0.01)


def test_rmat_degree_skew():
    g = G.synthetic("rmat", scale=14, edge_factor=8, seed=1)
    deg = g.degrees()
    assert deg.max() > 30 * deg.mean()  # heavy-tailed, unlike Erdos-Renyi


def _write_txt(path, n, edges, labels):
    with open(path, "w") as f:
        f.write("t # 0\n")
        for v, l in labels.items():
            f.write(f"v {v} {l}\n")
        for a, b, w in edges:
            f.write(f"e {a} {b} {w}\n")


def test_from_txt_sparse_ids_and_order(tmp_path):
    # ids are renumbered in ascending order; vertices first seen on an edge get label 0
    p = tmp_path / "g.txt"
    _write_txt(p, 0, [(40, 10, 2), (10, 7, 1), (40, 7, 0), (99, 40, 5)], {7: 3, 10: 1, 40: 2})
    g = G.from_txt(p)
    assert g.vlabels.tolist() == [3, 1, 2, 0]
    assert g.edges.tolist() == [[0, 1, 1], [0, 2, 0], [1, 0, 1], [1, 2, 2], [2, 0, 0],
                                [2, 1, 2], [2, 3, 5], [3, 2, 5]]


def test_host_convert_cli_matches_python(tmp_path):
    if not os.path.exists(CONVERT):
        pytest.skip("host CLIs not built")
    g = G.synthetic("er", n=300, m=800, n_vlabels=4, n_elabels=3, seed=8)
    G.to_txt(g, tmp_path / "g.txt")
    subprocess.run([CONVERT, str(tmp_path / "g.txt"), str(tmp_path / "g.bin")], check=True,
                   capture_output=True)
    h = G.CSRGraph.load(tmp_path / "g.bin")
    assert np.array_equal(g.edges, h.edges) and np.array_equal(g.vlabels, h.vlabels)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_converters_match_reference_convert(tmp_path, seed):
    ref = _ref()
    if ref is None:
        pytest.skip("reference build (oracle/_ref) not present")
    rng = np.random.default_rng(seed)
    # sparse, unsorted ids, repeated vertex lines, multi-labelled parallel edges
    ids = rng.choice(10000, size=40, replace=False)
    labels = {int(v): int(rng.integers(0, 5)) for v in ids}
    edges = []
    for _ in range(120):
        a, b = rng.choice(ids, size=2, replace=False)
        edges.append((int(a), int(b), int(rng.integers(0, 3))))
    _write_txt(tmp_path / "g.txt", 0, edges, labels)
    ref.ref_convert(str(tmp_path / "g.txt"), str(tmp_path / "ref.bin"))
    G.from_txt(tmp_path / "g.txt").save(tmp_path / "py.bin")
    assert (tmp_path / "py.bin").read_bytes() == (tmp_path / "ref.bin").read_bytes()
    if os.path.exists(CONVERT):
        subprocess.run([CONVERT, str(tmp_path / "g.txt"), str(tmp_path / "cli.bin")], check=True,
                       capture_output=True)
        assert (tmp_path / "cli.bin").read_bytes() == (tmp_path / "ref.bin").read_bytes()


def test_oracle_pattern_graph_matches_reference():
    ref = _ref()
    if ref is None:
        pytest.skip("reference build (oracle/_ref) not present")
    import oracle
    rng = np.random.default_rng(11)
    for _ in range(300):
        code = brute.random_code(rng, 6, 3, 3)
        a = oracle.pattern_graph(code)
        b = oracle.ref_pattern_graph(code)
        for x, y in zip(a, b):
            assert np.array_equal(x, y), code
