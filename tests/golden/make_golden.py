"""Writes the golden fixtures of this directory from the reference's own code.

Run in a checkout where oracle/_ref has been built from the reference sources
(`make -C oracle ref REF=/path/to/reference`):  python tests/golden/make_golden.py

* convert_<k>.txt / convert_<k>.bin  -- text inputs (sparse unsorted ids, vertices first
  named by an edge, repeated 'v' lines, parallel edges with different labels, a self-loop)
  and the reference `convert`'s binary output (Graph::from_txt + CSRGraph(Graph&) + save,
  graph.cc:38-214);
* prune_in.bin / prune_<min_sup>.bin  -- an input graph and the reference `prune_graph`
  (graph.cc:216-266) at several thresholds;
* pattern_graphs.json  -- DFS codes and the reference `CSRGraph(const DFSCode&)`
  (graph.cc:122-159) of each.
"""
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

import brute  # noqa: E402
import oracle  # noqa: E402
from gspan_amd import graph as G  # noqa: E402


def text_case(rng, n_ids, n_edges, self_loop):
    ids = rng.choice(100000, size=n_ids, replace=False) - 50000
    lines = ["t # 0"]
    for v in ids[: n_ids * 3 // 4]:  # a quarter of the ids appear only on edges
        lines.append(f"v {v} {rng.integers(-2, 5)}")
    for v in ids[:5]:  # repeated 'v' lines: the last one counts
        lines.append(f"v {v} {rng.integers(0, 9)}")
    for _ in range(n_edges):
        a, b = rng.choice(ids, size=2, replace=False)
        lines.append(f"e {a} {b} {rng.integers(0, 3)}")
        if rng.random() < 0.1:  # a parallel edge with another label
            lines.append(f"e {b} {a} {rng.integers(3, 5)}")
    if self_loop:
        lines.append(f"e {ids[0]} {ids[0]} 1")
    body = lines[1:]
    rng.shuffle(body)  # records in any order
    return "\n".join(lines[:1] + body) + "\n"


def main():
    if not oracle.ref_available():
        sys.exit("oracle/_ref is not built: run make -C oracle ref REF=<reference checkout>")
    rng = np.random.default_rng(20261203)
    for k, (ni, ne, sl) in enumerate([(30, 60, False), (200, 700, True), (1000, 4000, True)]):
        txt = os.path.join(HERE, f"convert_{k}.txt")
        with open(txt, "w") as f:
            f.write(text_case(rng, ni, ne, sl))
        oracle.ref_convert(txt, os.path.join(HERE, f"convert_{k}.bin"))
    g = G.synthetic("er", n=400, m=1200, n_vlabels=6, n_elabels=2, vlabel_zipf=1.3, seed=5)
    g.save(os.path.join(HERE, "prune_in.bin"))
    counts = np.bincount(g.vlabels)
    sups = sorted({1} | {int(c) for c in counts if c > 0} | {int(counts.max()) + 1})
    with tempfile.TemporaryDirectory() as td:
        for ms in sups:
            p = oracle.ref_prune(g, ms, td)
            p.save(os.path.join(HERE, f"prune_{ms}.bin"))
    cases = []
    for _ in range(120):
        code = brute.random_code(rng, 6, 3, 3)
        off, vl, ed = oracle.ref_pattern_graph(code)
        cases.append({"code": [list(map(int, e)) for e in code], "offsets": off.tolist(),
                      "vlabels": vl.tolist(), "edges": ed.tolist()})
    with open(os.path.join(HERE, "pattern_graphs.json"), "w") as f:
        json.dump(cases, f)
    print("golden fixtures written:", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
