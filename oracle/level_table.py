"""Per-level row counts of the benchmark configurations, from count-only passes on the CPU.

Planning tool beside the oracle (it drives oracle.level_estimate; nothing here is part of
the product).  For each configuration it prints one markdown table row per DFS level m:

* rows of canonical m-edge codes, any support -- what an embedding-list miner stores at
  level m when every pattern is frequent (an upper bound at the configuration's threshold):
  exact for m = 1 (CSR entries with label(from) <= label(to), gspan.cu:421) and m = 2
  (products of per-vertex label-neighbour counts, below), a Knuth random-probe estimate
  with its standard error for m >= 3;
* the largest single child list at m = 2, exact: for single-edge code (a, e, b) the
  forward child from v_1 with edge label e2 to label c has
  sum over vertices v with label b of N[v, (e, a)] * N[v, (e2, c)] rows (minus N[v, (e, a)]
  when (e2, c) = (e, a): the row's own edge), the child from v_0 the same over label-a
  vertices when it passes gspan.cu:475-477; N[v, (e, l)] = neighbours of v with edge label e
  and vertex label l;
* whether the level fits 2^31 - 1 rows (the reference's int prev_idx) and 288 GB of HBM at
  8 bytes per row.

    python oracle/level_table.py [cfg1 cfg2 ...] [--probes N]
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [HERE, os.path.dirname(HERE)]

import oracle  # noqa: E402
from gspan_amd.graph import synthetic  # noqa: E402

CONFIGS = {
    "cfg1": dict(kind="er", n=10000, m=40000, n_vlabels=5, vlabel_zipf=None, n_elabels=1,
                 min_sup=200, k=4),
    "cfg2": dict(kind="rmat", scale=20, edge_factor=8, n_vlabels=20, vlabel_zipf=1.1,
                 n_elabels=2, min_sup=2000, k=5),
    "cfg3": dict(kind="rmat", scale=20, edge_factor=8, n_vlabels=100, vlabel_zipf=1.0,
                 n_elabels=4, min_sup=500, k=5),
    "cfg4": dict(kind="rmat", scale=22, edge_factor=16, n_vlabels=50, vlabel_zipf=1.2,
                 n_elabels=4, min_sup=5000, k=6),
    "cfg5": dict(kind="rmat", scale=24, edge_factor=16, n_vlabels=200, vlabel_zipf=1.2,
                 n_elabels=8, min_sup=20000, k=6),
}
HBM = 288e9
INT_MAX = 2**31 - 1


def make(c: dict):
    if c["kind"] == "er":
        return synthetic("er", n=c["n"], m=c["m"], n_vlabels=c["n_vlabels"],
                         n_elabels=c["n_elabels"], seed=1)
    return synthetic("rmat", scale=c["scale"], edge_factor=c["edge_factor"],
                     n_vlabels=c["n_vlabels"], vlabel_zipf=c["vlabel_zipf"],
                     n_elabels=c["n_elabels"], seed=1)


def level2_exact(g):
    """(total rows of canonical 2-edge codes, largest canonical child list and its code)."""
    import scipy.sparse as sp
    n = g.nr_nodes
    e = g.edges
    vl = g.vlabels.astype(np.int64)
    VL = int(vl.max()) + 1
    EL = int(e[:, 2].max()) + 1 if e.shape[0] else 1
    C = EL * VL
    col = e[:, 2].astype(np.int64) * VL + vl[e[:, 1]]
    N = sp.csr_matrix((np.ones(e.shape[0]), (e[:, 0].astype(np.int64), col)), shape=(n, C))
    G, D = [], []
    for b in range(VL):
        Nb = N[vl == b]
        G.append((Nb.T @ Nb).toarray())
        D.append(np.asarray(Nb.sum(axis=0)).ravel())
    canon = {}

    def is_canon(code):
        if code not in canon:
            canon[code] = oracle.is_canonical(code) == 1
        return canon[code]

    total, best = 0, (0, None)
    for a in range(VL):
        for ee in range(EL):
            for b in range(a, VL):
                if D[b][ee * VL + a] == 0:
                    continue
                s0 = (0, 1, a, ee, b)
                for e2 in range(EL):
                    for c in range(a, VL):  # min_label = a (gspan.cu:455, 475)
                        x, y = ee * VL + a, e2 * VL + c
                        cnt = G[b][x, y] - (D[b][x] if x == y else 0)
                        kids = [((1, 2, -1, e2, c), cnt)]
                        if ee < e2 or (ee == e2 and b <= c):
                            x1 = ee * VL + b
                            kids.append(((0, 2, -1, e2, c),
                                         G[a][x1, y] - (D[a][x1] if x1 == y else 0)))
                        for edge, k in kids:
                            if k > 0 and is_canon((s0, edge)):
                                total += int(k)
                                if k > best[0]:
                                    best = (int(k), (s0, edge))
    return total, best


def fmt(x: float) -> str:
    return f"{x:.3g}" if x < 1e4 else f"{x:.2e}"


def run(name: str, c: dict, probes: int) -> list[str]:
    t0 = time.time()
    g = make(c)
    st = g.stats()
    tg = time.time() - t0
    k = c["k"]
    t0 = time.time()
    est = oracle.level_estimate(g, k, probes, seed=1)
    te = time.time() - t0
    t0 = time.time()
    l2_total, (l2_max, l2_code) = level2_exact(g) if k >= 2 else (0, (0, None))
    t2 = time.time() - t0
    out = [f"**{name}** ({'ER' if c['kind'] == 'er' else 'R-MAT scale %d' % c['scale']}: "
           f"n={st['n']:,}, nnz={st['nnz']:,}, max degree {st['max_degree']:,}, "
           f"{c['n_vlabels']} labels, {c['n_elabels']} edge labels, min_sup {c['min_sup']}, "
           f"k {k}; {probes:,} probes; graph {tg:.0f}s, estimate {te:.0f}s, level 2 {t2:.0f}s)",
           "",
           "| level m | rows of canonical m-edge codes | how | rows emitted into m | "
           "largest child list | > 2^31-1 rows | > 288 GB (8 B/row) |",
           "|---|---|---|---|---|---|---|"]
    for m in range(1, k + 1):
        rows, se = est["rows"][m], est["rows_se"][m]
        how = "exact" if m == 1 else f"estimate ± {fmt(se)}"
        if m == 2:
            rows, how = l2_total, f"exact (estimate {fmt(est['rows'][2])} ± {fmt(se)})"
        emit = "—" if m == 1 else f"{fmt(est['emitted'][m])} ± {fmt(est['emitted_se'][m])}"
        big = f"{l2_max:,} ({l2_code[1]})" if m == 2 and l2_code else ""
        out.append(f"| {m} | {fmt(rows)} | {how} | {emit} | {big} | "
                   f"{'yes' if rows > INT_MAX else 'no'} | {'yes' if rows * 8 > HBM else 'no'} |")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("configs", nargs="*", default=["cfg1", "cfg2", "cfg3", "cfg4"])
    ap.add_argument("--probes", type=int, default=200000)
    a = ap.parse_args()
    for name in a.configs:
        for line in run(name, CONFIGS[name], a.probes):
            print(line, flush=True)
        print(flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
