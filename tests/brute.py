"""Brute-force checkers that pin the oracle on small inputs.

Nothing here follows the reference's greedy rules: monomorphisms come from plain
backtracking, minimum DFS codes from exhaustive enumeration of every DFS code of a
pattern under the gSpan lexicographic order, frequent patterns from enumerating every
connected edge subset of the data graph.
"""
from __future__ import annotations

import itertools
from collections import defaultdict

import numpy as np


class SmallGraph:
    """Adjacency view of a CSR graph for brute force."""

    def __init__(self, g):
        self.n = g.nr_nodes
        self.vl = g.vlabels.tolist()
        self.entries = [tuple(e) for e in g.edges.tolist()]  # (from, to, label)
        self.adj = defaultdict(list)  # u -> [(entry id, to, label)]
        for i, (a, b, w) in enumerate(self.entries):
            self.adj[a].append((i, b, w))


def _code_vertex_labels(code):
    lab = {}
    for (a, b, fl, el, tl) in code:
        if fl != -1:
            lab[a] = fl
        if tl != -1:
            lab[b] = tl
    return lab


def monomorphisms(sg: SmallGraph, code):
    """All label-consistent monomorphisms of a DFS code, as tuples of CSR entry ids
    (entry j runs from the image of code[j].from to the image of code[j].to)."""
    lab = _code_vertex_labels(code)
    out = []
    m = len(code)

    def rec(j, phi, used_vertices, entries):
        if j == m:
            out.append(tuple(entries))
            return
        a, b, _fl, el, _tl = code[j]
        ua = phi[a]
        if b in phi:  # edge between two mapped vertices
            ub = phi[b]
            for (eid, to, w) in sg.adj[ua]:
                if to == ub and w == el and not _entry_used(sg, entries, eid):
                    rec(j + 1, phi, used_vertices, entries + [eid])
        else:
            for (eid, to, w) in sg.adj[ua]:
                if w != el or to in used_vertices or sg.vl[to] != lab[b]:
                    continue
                phi2 = dict(phi)
                phi2[b] = to
                rec(j + 1, phi2, used_vertices | {to}, entries + [eid])

    a0, b0, fl, el, tl = code[0]
    for u in range(sg.n):
        if sg.vl[u] != fl:
            continue
        for (eid, to, w) in sg.adj[u]:
            if w == el and sg.vl[to] == tl and to != u:
                rec(1, {a0: u, b0: to}, {u, to}, [eid])
    return out


def _entry_used(sg, entries, eid):
    a, b, w = sg.entries[eid]
    for x in entries:
        c, d, y = sg.entries[x]
        if y == w and ((c == a and d == b) or (c == b and d == a)):
            return True
    return False


def mni(sg: SmallGraph, code, omp_literal: bool = False) -> int:
    """Minimum-image support from brute-force monomorphisms.  With omp_literal, pattern
    vertex 0 only counts when some edge after the first leaves it."""
    embs = monomorphisms(sg, code)
    nv = max(max(e[0], e[1]) for e in code) + 1
    images = [set() for _ in range(nv)]
    for ent in embs:
        for j, eid in enumerate(ent):
            a, b, _ = sg.entries[eid]
            images[code[j][0]].add(a)
            images[code[j][1]].add(b)
    counts = [len(s) for s in images]
    if omp_literal and not any(e[0] == 0 for e in code[1:]):
        counts[0] = 0
    nz = [c for c in counts if c > 0]
    return min(nz) if nz else 0x7FFFFFFF


# ----------------------------------------------------------------- minimum DFS code

def _edge_key(pos, e):
    a, b, fl, el, tl = e
    if pos == 0:
        return (fl, el, tl)
    if a > b:  # backward: smaller target first, then edge label
        return (0, b, el, 0)
    return (1, -a, el, tl)  # forward: deeper source first, then (edge label, to label)


def code_key(code):
    return tuple(_edge_key(i, e) for i, e in enumerate(code))


def all_dfs_codes(vlabels, edges):
    """Every code produced by rightmost extension that covers all edges of a connected
    pattern graph.  vlabels: {vertex: label}; edges: [(u, v, label)] undirected."""
    m = len(edges)
    inc = defaultdict(list)
    for idx, (u, v, w) in enumerate(edges):
        inc[u].append((idx, v, w))
        inc[v].append((idx, u, w))
    codes = []

    def rec(code, phi_inv, order, used):
        # code: full-label DFS edges; phi_inv: code vertex -> graph vertex
        if len(code) == m:
            codes.append(tuple(code))
            return
        # rightmost path (code vertex ids)
        rm = []
        last = -1
        for i in range(len(code) - 1, -1, -1):
            a, b = code[i][0], code[i][1]
            if a < b and (not rm or b == last):
                rm.append(i)
                last = a
        v_n = code[rm[0]][1]
        rm_vertices = [code[i][0] for i in rm] + [v_n]
        g2c = {g: c for c, g in phi_inv.items()}
        # backward from v_n
        gn = phi_inv[v_n]
        for (idx, to, w) in inc[gn]:
            if idx in used or to not in g2c:
                continue
            cv = g2c[to]
            if cv in rm_vertices and cv != v_n:
                e = (v_n, cv, vlabels[gn], w, vlabels[to])
                rec(code + [e], phi_inv, order, used | {idx})
        # forward from rightmost path vertices
        nxt = len(phi_inv)
        for cv in rm_vertices:
            gv = phi_inv[cv]
            for (idx, to, w) in inc[gv]:
                if idx in used or to in g2c:
                    continue
                e = (cv, nxt, vlabels[gv], w, vlabels[to])
                p2 = dict(phi_inv)
                p2[nxt] = to
                rec(code + [e], p2, order, used | {idx})

    for idx, (u, v, w) in enumerate(edges):
        for (a, b) in ((u, v), (v, u)):
            rec([(0, 1, vlabels[a], w, vlabels[b])], {0: a, 1: b}, None, {idx})
    return codes


def to_placeholder(code):
    """Full-label code -> the reference's form (-1 for labels it does not store)."""
    out = []
    for i, (a, b, fl, el, tl) in enumerate(code):
        if i == 0:
            out.append((a, b, fl, el, tl))
        elif a > b:
            out.append((a, b, -1, el, -1))
        else:
            out.append((a, b, -1, el, tl))
    return tuple(out)


def pattern_of_code(code):
    """(vertex labels dict, undirected edge list) of a placeholder-form code."""
    lab = _code_vertex_labels(code)
    edges = [(a, b, el) for (a, b, _fl, el, _tl) in code]
    return lab, edges


def min_dfs_code(vlabels, edges):
    codes = all_dfs_codes(vlabels, edges)
    return to_placeholder(min(codes, key=code_key))


def is_min_code(code) -> bool:
    lab, edges = pattern_of_code(code)
    return min_dfs_code(lab, edges) == tuple(tuple(e) for e in code)


# ----------------------------------------------------------------- frequent patterns

def connected_edge_subsets(sg: SmallGraph, kmax: int):
    """All connected sets of undirected data edges with 1..kmax edges."""
    und = {}
    for i, (a, b, w) in enumerate(sg.entries):
        if a < b:
            und[(a, b, w)] = len(und)
    elist = list(und)
    inc = defaultdict(list)
    for idx, (a, b, w) in enumerate(elist):
        inc[a].append(idx)
        inc[b].append(idx)
    seen = set()
    frontier = [frozenset([i]) for i in range(len(elist))]
    seen.update(frontier)
    result = list(frontier)
    for _size in range(2, kmax + 1):
        nxt = []
        for s in frontier:
            verts = set()
            for i in s:
                verts.add(elist[i][0])
                verts.add(elist[i][1])
            for v in verts:
                for j in inc[v]:
                    if j not in s:
                        t = s | {j}
                        if t not in seen:
                            seen.add(t)
                            nxt.append(t)
        result.extend(nxt)
        frontier = nxt
    return elist, result


def frequent_patterns(g, k: int, min_sup: int, omp_literal: bool = False):
    """{min DFS code (placeholder form): MNI} for every connected pattern with at most k
    edges whose MNI (over all of its monomorphisms) is at least min_sup."""
    sg = SmallGraph(g)
    elist, subsets = connected_edge_subsets(sg, k)
    patterns = set()
    for s in subsets:
        verts = sorted({x for i in s for x in elist[i][:2]})
        lab = {v: sg.vl[v] for v in verts}
        edges = [elist[i] for i in s]
        patterns.add(min_dfs_code(lab, edges))
    out = {}
    for code in patterns:
        sup = mni(sg, code, omp_literal=omp_literal)
        if sup >= min_sup:
            out[code] = sup
    return out


def random_small_graph(rng: np.random.Generator, n: int, m: int, n_vlabels: int,
                       n_elabels: int):
    from gspan_amd.graph import csr_from_undirected
    pairs = set()
    while len(pairs) < m:
        a, b = rng.integers(0, n, size=2)
        if a != b:
            pairs.add((min(a, b), max(a, b)))
    pairs = sorted(pairs)
    u = np.array([p[0] for p in pairs])
    v = np.array([p[1] for p in pairs])
    el = rng.integers(0, n_elabels, size=len(pairs))
    vl = rng.integers(0, n_vlabels, size=n)
    return csr_from_undirected(n, u, v, el, vl)


def random_code(rng: np.random.Generator, max_edges: int, n_vlabels: int, n_elabels: int):
    """A random DFS code (canonical or not) from a random DFS traversal of a random
    small connected labelled graph."""
    while True:
        nv = int(rng.integers(2, max_edges + 2))
        edges = {}
        # random spanning tree then extra edges
        for v in range(1, nv):
            u = int(rng.integers(0, v))
            edges[(u, v)] = int(rng.integers(0, n_elabels))
        extra = int(rng.integers(0, max_edges - (nv - 1) + 1)) if max_edges > nv - 1 else 0
        tries = 0
        while extra > 0 and tries < 50:
            tries += 1
            a, b = sorted(int(x) for x in rng.integers(0, nv, size=2))
            if a != b and (a, b) not in edges:
                edges[(a, b)] = int(rng.integers(0, n_elabels))
                extra -= 1
        if len(edges) > max_edges:
            continue
        lab = {v: int(rng.integers(0, n_vlabels)) for v in range(nv)}
        elist = [(a, b, w) for (a, b), w in edges.items()]
        codes = all_dfs_codes(lab, elist)
        return to_placeholder(codes[int(rng.integers(0, len(codes)))])
