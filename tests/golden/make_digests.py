"""Writes tests/golden/report_digests.json: SHA-256 digests of the oracle's report lines
(gspan.cu:67-77 format, in the reference's order) with its counters, for mining runs too
slow to repeat inside a GPU test.  Run in this repository with the oracle built:

    python tests/golden/make_digests.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

import oracle  # noqa: E402
from gspan_amd.graph import synthetic  # noqa: E402

CASES = {
    # R-MAT scale 12, 100 labels Zipf(1.0), 4 edge labels (configuration 3's distributions),
    # k = 3, min_sup 4: the last level has more than 2^21 live child sets
    "rmat12_cfg3_k3_ms4": dict(graph=dict(kind="rmat", scale=12, edge_factor=8, n_vlabels=100,
                                          vlabel_zipf=1.0, n_elabels=4, seed=1), k=3, min_sup=4),
}


def digest(patterns) -> str:
    return hashlib.sha256(oracle.format_report(patterns).encode()).hexdigest()


def main() -> int:
    out = {}
    for name, c in CASES.items():
        g = synthetic(**c["graph"])
        pats, st = oracle.fsm(g, c["k"], c["min_sup"])
        out[name] = dict(c, patterns=len(pats), extended=st["extended"],
                         canonical=st["canonical"], candidates=st["candidates"],
                         sha256=digest(pats))
        print(name, out[name], flush=True)
    with open(os.path.join(HERE, "report_digests.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
