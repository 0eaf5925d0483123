"""Text ingestion on the GPU (gs_text_parse: Graph::from_txt + CSRGraph(Graph&),
graph.cc:38-112) against the host restatement `graph.from_txt`, which tests/test_graph_io.py
pins byte for byte against the reference's own `convert` built from its sources."""
import numpy as np
import pytest

from gspan_amd import graph as G

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def same(a, b):
    assert a.nr_nodes == b.nr_nodes and a.nr_edges == b.nr_edges
    np.testing.assert_array_equal(a.offsets, b.offsets)
    np.testing.assert_array_equal(a.vlabels, b.vlabels)
    np.testing.assert_array_equal(a.edges, b.edges)


HAND = {
    # ids named only by edges get label 0; the last 'v' record wins; negative ids and
    # labels; a self-loop (two entries) and a multi-edge; blank, comment and 't' lines;
    # CRLF endings; no newline at the end
    "mixed": "t # 0\r\nv 5 7\r\nv -3 -2\r\n\r\ne 5 9 1\r\ne 9 -3 -4\r\n# comment\r\n"
             "e 5 5 2\r\ne 9 5 1\r\ne 9 5 0\r\nv 5 8\r\nx 1 2\r\ne 100 -3 6",
    "edges_only": "e 3 1 0\ne 1 2 0\ne 2 3 1\n",
    "vertices_only": "v 2 1\nv 0 3\nv 1 1\n",
    "empty": "",
    "blank": "\n\n  \n",
    "spaces": "  v   1   2  \n\te 1 0 3\t\n v 0 -1",
}

BAD = {
    "e 1 2\n": 0,                   # missing label: the reference would read the next line
    "v 1 2\nv 1 2 3\n": 1,          # extra token
    "v 1 2\ne 1 2 x\n": 1,          # non-integer argument
    "t # 0\nfoo e 1 2 3\n": 1,      # op letter after an unknown op
    "v 1 99999999999\n": 0,         # beyond int32
}


def test_text_ingest_hand_cases(tmp_path):
    for name, text in HAND.items():
        _announce(f"ingest hand case {name}")
        p = tmp_path / f"{name}.txt"
        p.write_bytes(text.encode())
        same(G.from_txt_gpu(p), G.from_txt(p))


def test_text_ingest_rejects_cross_line_records(tmp_path):
    for i, (text, line) in enumerate(BAD.items()):
        _announce(f"ingest malformed case {i}")
        p = tmp_path / f"bad{i}.txt"
        p.write_bytes(text.encode())
        with pytest.raises(ValueError, match=f"line {line + 1} "):
            G.from_txt_gpu(p)


def test_text_ingest_synthetic_graphs(tmp_path):
    for kind, kw in [("er", dict(n=3000, m=12000, n_vlabels=5, n_elabels=3, seed=2)),
                     ("rmat", dict(scale=14, edge_factor=8, n_vlabels=20, vlabel_zipf=1.1,
                                   n_elabels=2, seed=3))]:
        _announce(f"ingest synthetic {kind}")
        g = G.synthetic(kind, **kw)
        p = tmp_path / f"{kind}.txt"
        G.to_txt(g, p)
        gg = G.from_txt_gpu(p)
        same(gg, G.from_txt(p))
        same(gg, g)  # the writer's graph comes back unchanged


def test_text_ingest_shuffled_records(tmp_path):
    """Records in any order, ids sparse and unsorted: same CSR as the host reader."""
    rng = np.random.default_rng(11)
    ids = rng.choice(10**6, size=500, replace=False) - 500_000
    lines = [f"v {i} {rng.integers(-3, 4)}" for i in ids]
    for _ in range(3000):
        a, b = rng.choice(ids, 2)
        lines.append(f"e {a} {b} {rng.integers(0, 3)}")
    rng.shuffle(lines)
    p = tmp_path / "shuf.txt"
    p.write_text("\n".join(lines) + "\n")
    _announce("ingest shuffled records")
    same(G.from_txt_gpu(p), G.from_txt(p))


def test_text_ingest_matches_reference_golden(tmp_path):
    """Byte-identical to the reference's own convert on the committed fixtures."""
    import glob
    import os
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    cases = sorted(glob.glob(os.path.join(gold, "convert_*.txt")))
    assert cases
    for txt in cases:
        _announce(f"ingest golden {os.path.basename(txt)}")
        G.from_txt_gpu(txt).save(tmp_path / "gpu.bin")
        assert (tmp_path / "gpu.bin").read_bytes() == open(txt[:-4] + ".bin", "rb").read()
