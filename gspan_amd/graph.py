"""Data-graph types, the reference's file formats, and seeded synthetic generators.

The binary CSR format and the CSR construction order are kept identical to the
reference so that files move between the two unchanged:

* binary file (graph.cc:178-214): native-endian int32 ``n, nnz, offsets[n+1],
  vlabels[n]`` followed by ``nnz`` 12-byte ``Edge{from, to, label}`` records;
* text format (graph.cc:38-64): tokens ``t # id`` / ``v id label`` / ``e u v label``,
  every edge stored in both directions (graph.cc:32-36);
* CSR build (graph.cc:66-112): vertices renumbered in ascending id order, each
  adjacency sorted by ``(to, label)``.
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

__all__ = [
    "CSRGraph",
    "from_txt",
    "csr_from_undirected",
    "erdos_renyi",
    "rmat",
    "zipf_labels",
    "uniform_labels",
    "synthetic",
]


@dataclass
class CSRGraph:
    """CSR data graph, graph.h:51-65 (``nr_nodes``, ``nr_edges`` = directed entries)."""

    offsets: np.ndarray  # int32 [n+1]
    vlabels: np.ndarray  # int32 [n]
    edges: np.ndarray  # int32 [nnz, 3] = (from, to, label)

    def __post_init__(self):
        self.offsets = np.ascontiguousarray(self.offsets, dtype=np.int32)
        self.vlabels = np.ascontiguousarray(self.vlabels, dtype=np.int32)
        self.edges = np.ascontiguousarray(self.edges, dtype=np.int32).reshape(-1, 3)
        if self.offsets.shape[0] != self.vlabels.shape[0] + 1:
            raise ValueError("offsets must have n+1 entries")
        if int(self.offsets[-1]) != self.edges.shape[0]:
            raise ValueError("offsets[n] must equal the number of edge entries")

    @property
    def nr_nodes(self) -> int:
        return int(self.vlabels.shape[0])

    @property
    def nr_edges(self) -> int:
        return int(self.edges.shape[0])

    # graph.cc:178-200
    @classmethod
    def load(cls, path: str | os.PathLike) -> "CSRGraph":
        with open(path, "rb") as f:
            head = np.fromfile(f, dtype=np.int32, count=2)
            if head.shape[0] < 1:
                raise RuntimeError("nr_nodes required")
            if head.shape[0] < 2:
                raise RuntimeError("nr_edges required")
            n, nnz = int(head[0]), int(head[1])
            if n < 0 or nnz < 0:
                raise RuntimeError("bad header")
            offsets = np.fromfile(f, dtype=np.int32, count=n + 1)
            if offsets.shape[0] != n + 1:
                raise RuntimeError("bad offsets")
            vlabels = np.fromfile(f, dtype=np.int32, count=n)
            if vlabels.shape[0] != n:
                raise RuntimeError("bad vertex labels")
            edges = np.fromfile(f, dtype=np.int32, count=3 * nnz)
            if edges.shape[0] != 3 * nnz:
                raise RuntimeError("bad edges")
        return cls(offsets, vlabels, edges.reshape(nnz, 3))

    # graph.cc:202-214
    def save(self, path: str | os.PathLike) -> None:
        with open(path, "wb") as f:
            np.array([self.nr_nodes, self.nr_edges], dtype=np.int32).tofile(f)
            self.offsets.tofile(f)
            self.vlabels.tofile(f)
            self.edges.tofile(f)

    def degrees(self) -> np.ndarray:
        return np.diff(self.offsets)

    def stats(self) -> dict:
        deg = self.degrees().astype(np.int64)
        return {
            "n": self.nr_nodes,
            "nnz": self.nr_edges,
            "undirected_edges": self.nr_edges // 2,
            "max_degree": int(deg.max()) if deg.size else 0,
            "sum_deg_sq": int((deg * deg).sum()),
            "vertex_labels": int(np.unique(self.vlabels).shape[0]),
            "edge_labels": int(np.unique(self.edges[:, 2]).shape[0]) if self.nr_edges else 0,
        }


def _csr_from_directed(n: int, src: np.ndarray, dst: np.ndarray, lab: np.ndarray,
                       vlabels: np.ndarray) -> CSRGraph:
    """Directed entries -> CSR with every adjacency sorted by (to, label) (graph.cc:96-100)."""
    order = np.lexsort((lab, dst, src))
    src, dst, lab = src[order], dst[order], lab[order]
    counts = np.bincount(src, minlength=n).astype(np.int64)
    offsets = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=offsets[1:])
    if offsets[-1] > np.iinfo(np.int32).max:
        raise ValueError("graph too large for int32 CSR offsets")
    edges = np.stack([src, dst, lab], axis=1).astype(np.int32)
    return CSRGraph(offsets.astype(np.int32), vlabels.astype(np.int32), edges)


def csr_from_undirected(n: int, u: np.ndarray, v: np.ndarray, elabels: np.ndarray,
                        vlabels: np.ndarray) -> CSRGraph:
    """Undirected edge list -> CSR holding each edge in both directions (graph.cc:32-36)."""
    u = np.asarray(u, dtype=np.int64)
    v = np.asarray(v, dtype=np.int64)
    el = np.asarray(elabels, dtype=np.int64)
    src = np.concatenate([u, v])
    dst = np.concatenate([v, u])
    lab = np.concatenate([el, el])
    return _csr_from_directed(n, src, dst, lab, np.asarray(vlabels))


def from_txt(path: str | os.PathLike, verbose: bool = False) -> CSRGraph:
    """Graph::from_txt + CSRGraph(Graph&) (graph.cc:38-112).

    Tokens are read one at a time; ``t`` consumes two more tokens, ``v`` two ints, ``e``
    three ints, anything else is skipped.  Vertices are created on first mention (label
    0 until a ``v`` line sets it), renumbered in ascending id order.
    """
    with open(path, "r") as f:
        toks = f.read().split()
    vlab: dict[int, int] = {}
    eu: list[int] = []
    ev: list[int] = []
    el: list[int] = []
    i, nt = 0, len(toks)
    while i < nt:
        op = toks[i]
        i += 1
        c = op[0]
        if c == "t":
            i += 2
        elif c == "v":
            if i + 2 > nt:
                break
            vid, lab = int(toks[i]), int(toks[i + 1])
            i += 2
            vlab[vid] = lab
        elif c == "e":
            if i + 3 > nt:
                break
            a, b, w = int(toks[i]), int(toks[i + 1]), int(toks[i + 2])
            i += 3
            vlab.setdefault(a, 0)
            vlab.setdefault(b, 0)
            eu.append(a)
            ev.append(b)
            el.append(w)
    ids = np.array(sorted(vlab), dtype=np.int64)
    n = ids.shape[0]
    if verbose:
        print(f"|V| = {n} |E| = {len(eu)}")
    labels = np.array([vlab[int(x)] for x in ids], dtype=np.int64)
    if len(eu) == 0:
        return CSRGraph(np.zeros(n + 1, np.int32), labels, np.zeros((0, 3), np.int32))
    u = np.searchsorted(ids, np.array(eu, dtype=np.int64))
    v = np.searchsorted(ids, np.array(ev, dtype=np.int64))
    return csr_from_undirected(n, u, v, np.array(el, dtype=np.int64), labels)


def from_txt_gpu(path: str | os.PathLike, device: int = 0, verbose: bool = False) -> CSRGraph:
    """Graph::from_txt + CSRGraph(Graph&) (graph.cc:38-112) on the GPU (gs_text_parse).

    The file must hold one record per line, as gSpan files do; a line the reference would
    read across a line boundary raises ValueError naming the line.
    """
    import ctypes

    from . import _lib as L
    with open(path, "rb") as f:
        data = f.read()
    lib = L.lib()
    h = ctypes.c_void_p()
    n, nnz, bad = ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong(-1)
    rc = lib.gs_text_parse(data, len(data), device, ctypes.byref(h), ctypes.byref(n),
                           ctypes.byref(nnz), ctypes.byref(bad))
    if rc == L.GS_ERR_INVALID and bad.value >= 0:
        raise ValueError(f"{path}: line {bad.value + 1} is not one whole record")
    if rc != L.GS_OK:
        raise L.GsError(rc, "gs_text_parse failed")
    try:
        off = np.zeros(n.value + 1, np.int32)
        vl = np.zeros(max(n.value, 1), np.int32)
        ed = np.zeros((max(nnz.value, 1), 3), np.int32)
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        rc = lib.gs_text_copy(h, ip(off), ip(vl), ip(ed))
        if rc != L.GS_OK:
            raise L.GsError(rc, "gs_text_copy failed")
    finally:
        lib.gs_text_free(h)
    if verbose:
        print(f"|V| = {n.value} |E| = {nnz.value // 2}")
    return CSRGraph(off, vl[:n.value], ed[:nnz.value])


def to_txt(g: CSRGraph, path: str | os.PathLike) -> None:
    """Writes the text format read by graph.cc:38-64 (each undirected edge once)."""
    e = g.edges
    keep = e[:, 0] < e[:, 1]
    with open(path, "w") as f:
        f.write("t # 0\n")
        for vid, lab in enumerate(g.vlabels.tolist()):
            f.write(f"v {vid} {lab}\n")
        for a, b, w in e[keep].tolist():
            f.write(f"e {a} {b} {w}\n")


# ------------------------------------------------------------------ generators

def uniform_labels(rng: np.random.Generator, count: int, n_labels: int) -> np.ndarray:
    return rng.integers(0, n_labels, size=count, dtype=np.int64)


def zipf_labels(rng: np.random.Generator, count: int, n_labels: int, s: float) -> np.ndarray:
    """i.i.d. labels with P(label = i) proportional to 1 / (i+1)^s, i in [0, n_labels)."""
    w = 1.0 / np.arange(1, n_labels + 1, dtype=np.float64) ** s
    cdf = np.cumsum(w / w.sum())
    cdf[-1] = 1.0
    return np.searchsorted(cdf, rng.random(count), side="right").astype(np.int64)


def _dedup_pairs(u: np.ndarray, v: np.ndarray):
    keep = u != v
    u, v = u[keep], v[keep]
    lo = np.minimum(u, v)
    hi = np.maximum(u, v)
    key = np.unique((lo << 32) | hi)
    return key >> 32, key & 0xFFFFFFFF


def erdos_renyi(n: int, m: int, rng: np.random.Generator):
    """m distinct unordered pairs drawn uniformly, no self-loops (sorted by (lo, hi))."""
    if m > n * (n - 1) // 2:
        raise ValueError("too many edges")
    keys = np.zeros(0, dtype=np.int64)
    while keys.shape[0] < m:
        need = m - keys.shape[0]
        a = rng.integers(0, n, size=2 * need + 16, dtype=np.int64)
        b = rng.integers(0, n, size=2 * need + 16, dtype=np.int64)
        keep = a != b
        a, b = a[keep], b[keep]
        k = (np.minimum(a, b) << 32) | np.maximum(a, b)
        # keep first occurrences in draw order so the edge set depends only on the seed
        _, first = np.unique(k, return_index=True)
        k = k[np.sort(first)]
        k = k[~np.isin(k, keys)]
        keys = np.concatenate([keys, k[:need]])
    keys = np.sort(keys)
    return keys >> 32, keys & 0xFFFFFFFF


def rmat(scale: int, edge_factor: int, rng: np.random.Generator,
         a: float = 0.57, b: float = 0.19, c: float = 0.19, d: float = 0.05):
    """R-MAT edge pairs (recursive quadrant choice per bit), self-loops removed,
    unordered pairs deduplicated.  Returns (lo, hi) sorted."""
    total = edge_factor << scale
    u = np.zeros(total, dtype=np.int64)
    v = np.zeros(total, dtype=np.int64)
    t_ab = a + b
    t_abc = a + b + c
    for bit in range(scale):
        r = rng.random(total)
        ubit = r >= t_ab  # quadrants c, d
        vbit = ((r >= a) & (r < t_ab)) | (r >= t_abc)  # quadrants b, d
        u |= ubit.astype(np.int64) << bit
        v |= vbit.astype(np.int64) << bit
        del r, ubit, vbit
    return _dedup_pairs(u, v)


def synthetic(kind: str, *, n: int = 0, m: int = 0, scale: int = 0, edge_factor: int = 8,
              n_vlabels: int = 1, vlabel_zipf: float | None = None, n_elabels: int = 1,
              seed: int = 1) -> CSRGraph:
    """Seeded synthetic labelled graph in the reference's CSR layout.

    ``kind="er"``: ``m`` uniform distinct pairs over ``n`` vertices.
    ``kind="rmat"``: R-MAT with ``2**scale`` vertices and ``edge_factor * 2**scale`` draws.
    Vertex labels: Zipf(``vlabel_zipf``) over ``n_vlabels`` when given, else uniform.
    Edge labels: uniform over ``n_elabels``.
    """
    rng = np.random.default_rng(seed)
    if kind == "er":
        lo, hi = erdos_renyi(n, m, rng)
        nv = n
    elif kind == "rmat":
        lo, hi = rmat(scale, edge_factor, rng)
        nv = 1 << scale
    else:
        raise ValueError(f"unknown graph kind {kind!r}")
    if vlabel_zipf is None:
        vl = uniform_labels(rng, nv, n_vlabels)
    else:
        vl = zipf_labels(rng, nv, n_vlabels, vlabel_zipf)
    el = uniform_labels(rng, lo.shape[0], n_elabels)
    return csr_from_undirected(nv, lo, hi, el, vl)
