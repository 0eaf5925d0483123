"""ctypes front end of the CPU oracle (oracle/gspan_oracle.cc).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` as the checker / timed CPU baseline, never by the
product package.

``prune_graph`` is pinned against the reference's own prune_graph (oracle/_ref, and its
outputs in tests/golden) and a loop restatement; ``format_report`` order and format are
parity unpinned (restated from gspan.cu:67-77, 1142-1168).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Callable

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None

REPORT_CB = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                             ctypes.c_int, ctypes.c_void_p)
CHILD_CB = ctypes.CFUNCTYPE(None, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                            ctypes.c_longlong, ctypes.c_void_p)

MNI_GPU = 0
MNI_OMP_LITERAL = 1


def build(force: bool = False) -> None:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE, "all"], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        c_int, c_ll = ctypes.c_int, ctypes.c_longlong
        L.oracle_fsm.argtypes = [P(c_int), P(c_int), P(c_int), c_int, c_int, c_int, c_int, c_int,
                                 c_int, REPORT_CB, ctypes.c_void_p, P(c_ll)]
        L.oracle_fsm.restype = c_int
        L.oracle_fsm_bounded.argtypes = L.oracle_fsm.argtypes + [ctypes.c_double]
        L.oracle_fsm_bounded.restype = c_int
        L.oracle_is_canonical.argtypes = [P(c_int), c_int]
        L.oracle_is_canonical.restype = c_int
        L.oracle_support.argtypes = [P(c_int), P(c_int), P(c_int), c_int, c_int, P(c_int), c_int,
                                     P(P(c_int)), P(c_ll), c_int, c_int]
        L.oracle_support.restype = c_int
        L.oracle_extend.argtypes = [P(c_int), P(c_int), P(c_int), c_int, c_int, P(c_int), c_int,
                                    P(P(c_int)), P(c_ll), c_int, CHILD_CB, ctypes.c_void_p]
        L.oracle_extend.restype = c_ll
        L.oracle_seed.argtypes = [P(c_int), P(c_int), P(c_int), c_int, c_int, CHILD_CB,
                                  ctypes.c_void_p]
        L.oracle_seed.restype = c_int
        L.oracle_pattern_graph.argtypes = [P(c_int), c_int, P(c_int), P(c_int), P(c_int),
                                           P(c_int)]
        L.oracle_pattern_graph.restype = c_int
        L.oracle_level_counts.argtypes = [P(c_int), P(c_int), P(c_int), c_int, c_int, c_int,
                                          c_int, P(c_ll), P(c_ll)]
        L.oracle_level_counts.restype = c_int
        c_d = ctypes.c_double
        L.oracle_level_estimate.argtypes = [P(c_int), P(c_int), P(c_int), c_int, c_int, c_int,
                                            c_ll, ctypes.c_ulonglong, c_int, P(c_d), P(c_d),
                                            P(c_d), P(c_d)]
        L.oracle_level_estimate.restype = c_int
        _lib = L
    return _lib


def _ip(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _graph_args(g):
    return (_ip(g.offsets), _ip(g.vlabels), _ip(g.edges), g.nr_nodes, g.nr_edges)


def _code_arr(code) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(code, dtype=np.int32).reshape(-1, 5))


def fsm(g, k: int, min_sup: int, mni_mode: int = MNI_GPU, nthreads: int = 0,
        on_report: Callable | None = None, max_seconds: float = 0.0):
    """Full mining run.  Returns (patterns, stats); patterns = [(code tuple, support)]
    in the reference's report order.  With max_seconds > 0 the run stops at the first
    pattern visited after that time (stats["truncated"] = 1)."""
    out = []

    def cb(cp, m, sup, _u):
        code = tuple(tuple(cp[5 * i + j] for j in range(5)) for i in range(m))
        out.append((code, sup))
        if on_report is not None:
            on_report(code, sup)
        return 0

    st = (ctypes.c_longlong * 8)()
    rc = lib().oracle_fsm_bounded(*_graph_args(g), k, min_sup, mni_mode, nthreads, REPORT_CB(cb),
                                  None, st, float(max_seconds))
    if rc not in (0, 2):
        raise RuntimeError("oracle_fsm failed")
    names = ["extended", "candidates", "canonical", "reported", "overruns", "t_extend_us",
             "t_support_us", "t_canon_us"]
    d = dict(zip(names, list(st)))
    d["truncated"] = int(rc == 2)
    return out, d


def is_canonical(code) -> int:
    c = _code_arr(code)
    return lib().oracle_is_canonical(_ip(c), c.shape[0])


def _layers(layers):
    arrs = [np.ascontiguousarray(np.asarray(x, dtype=np.int32).reshape(-1, 2)) for x in layers]
    ptrs = (ctypes.POINTER(ctypes.c_int) * len(arrs))(*[_ip(a) for a in arrs])
    counts = (ctypes.c_longlong * len(arrs))(*[a.shape[0] for a in arrs])
    return arrs, ptrs, counts


def support(g, code, layers, mni_mode: int = MNI_GPU, nthreads: int = 0) -> int:
    c = _code_arr(code)
    arrs, ptrs, counts = _layers(layers)
    return lib().oracle_support(*_graph_args(g), _ip(c), c.shape[0], ptrs, counts, mni_mode,
                                nthreads)


def extend(g, code, layers, nthreads: int = 0):
    """One extension step; returns [(child_edge5 tuple, rows int32[count, 2])] in the
    reference's recursion order."""
    c = _code_arr(code)
    arrs, ptrs, counts = _layers(layers)
    out = []

    def cb(ep, rp, cnt, _u):
        e5 = tuple(ep[j] for j in range(5))
        rows = np.ctypeslib.as_array(rp, shape=(cnt * 2,)).reshape(cnt, 2).copy() if cnt else \
            np.zeros((0, 2), np.int32)
        out.append((e5, rows))

    total = lib().oracle_extend(*_graph_args(g), _ip(c), c.shape[0], ptrs, counts, nthreads,
                                CHILD_CB(cb), None)
    if total < 0:
        raise RuntimeError("oracle_extend failed")
    return out


def seed(g):
    """Single-edge codes and their rows: [(edge5, rows int32[count, 2])]."""
    out = []

    def cb(ep, rp, cnt, _u):
        e5 = tuple(ep[j] for j in range(5))
        rows = np.ctypeslib.as_array(rp, shape=(cnt * 2,)).reshape(cnt, 2).copy()
        out.append((e5, rows))

    lib().oracle_seed(*_graph_args(g), CHILD_CB(cb), None)
    return out


def level_counts(g, k: int, nthreads: int = 0):
    """Exact level sizes of a full mining run at min_support 1: (rows, emitted), where
    rows[m] = rows of the canonical m-edge codes and emitted[m] = child rows extension
    emits into level m (index 0 unused)."""
    rows = (ctypes.c_longlong * 9)()
    emitted = (ctypes.c_longlong * 9)()
    if lib().oracle_level_counts(*_graph_args(g), k, nthreads, rows, emitted) != 0:
        raise RuntimeError("oracle_level_counts failed")
    return list(rows)[:k + 1], list(emitted)[:k + 1]


def level_estimate(g, k: int, probes: int, seed: int = 1, nthreads: int = 0):
    """Knuth random-probe estimates of the same level sizes (oracle_level_estimate):
    dict of lists rows, rows_se, emitted, emitted_se (standard errors), index 0 unused."""
    arr = [(ctypes.c_double * 9)() for _ in range(4)]
    if lib().oracle_level_estimate(*_graph_args(g), k, probes, seed, nthreads, *arr) != 0:
        raise RuntimeError("oracle_level_estimate failed")
    r, r2, e, e2 = ([x for x in a][:k + 1] for a in arr)
    se = lambda m1, m2: [max(b - a * a, 0.0) ** 0.5 / probes ** 0.5 for a, b in zip(m1, m2)]
    return {"rows": r, "rows_se": se(r, r2), "emitted": e, "emitted_se": se(e, e2)}


def pattern_graph(code):
    c = _code_arr(code)
    dims = (ctypes.c_int * 2)()
    lib().oracle_pattern_graph(_ip(c), c.shape[0], dims, None, None, None)
    n, nnz = dims[0], dims[1]
    off = np.zeros(n + 1, np.int32)
    vl = np.zeros(n, np.int32)
    ed = np.zeros((nnz, 3), np.int32)
    lib().oracle_pattern_graph(_ip(c), c.shape[0], dims, _ip(off), _ip(vl), _ip(ed))
    return off, vl, ed


def prune_graph(g):
    """Restatement of prune_graph (graph.cc:216-266): vertices whose label occurs fewer than
    min_sup times are dropped, the others renumbered in ascending id order; each kept
    vertex keeps its edges to kept vertices in CSR order.  Returns a function of min_sup."""
    def run(min_sup: int):
        from gspan_amd.graph import CSRGraph
        counts = np.bincount(g.vlabels - g.vlabels.min()) if g.nr_nodes else np.zeros(0, np.int64)
        keep = counts[g.vlabels - g.vlabels.min()] >= min_sup if g.nr_nodes else np.zeros(0, bool)
        newid = np.cumsum(keep) - 1
        e = g.edges
        own = np.repeat(np.arange(g.nr_nodes), np.diff(g.offsets))
        ke = keep[own] & keep[e[:, 1]] if e.shape[0] else np.zeros(0, bool)
        out = np.stack([newid[e[ke, 0]], newid[e[ke, 1]], e[ke, 2]], axis=1).astype(np.int32)
        kv = np.nonzero(keep)[0]
        deg = np.bincount(newid[own[ke]], minlength=kv.shape[0]) if kv.shape[0] else np.zeros(0)
        off = np.zeros(kv.shape[0] + 1, np.int64)
        np.cumsum(deg, out=off[1:])
        return CSRGraph(off.astype(np.int32), g.vlabels[kv], out.reshape(-1, 3))
    return run


def ref_prune(g, min_sup: int, tmpdir: str):
    """The reference's own prune_graph (oracle/_ref) on g; returns the pruned CSRGraph."""
    from gspan_amd.graph import CSRGraph
    src = os.path.join(tmpdir, "in.bin")
    dst = os.path.join(tmpdir, "out.bin")
    g.save(src)
    ref_lib().ref_prune_file(src.encode(), min_sup, dst.encode())
    return CSRGraph.load(dst)


def format_report(patterns) -> str:
    """The reference's stdout lines (gspan.cu:67-77)."""
    lines = []
    for code, sup in patterns:
        lines.append("".join("(%d %d %d %d %d)" % e for e in code) + ": %d" % sup)
    return "\n".join(lines) + ("\n" if lines else "")


# ---------------------------------------------------------------- reference build (_ref)

def ref_available() -> bool:
    return os.path.exists(os.path.join(REF_DIR, "convert")) and \
        os.path.exists(os.path.join(REF_DIR, "libref_graph.so"))


_ref_lib = None


def ref_lib():
    global _ref_lib
    if _ref_lib is None:
        L = ctypes.CDLL(os.path.join(REF_DIR, "libref_graph.so"))
        P = ctypes.POINTER
        L.ref_pattern_graph.argtypes = [P(ctypes.c_int), ctypes.c_int, P(ctypes.c_int),
                                        P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int)]
        L.ref_prune_file.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p]
        _ref_lib = L
    return _ref_lib


def ref_pattern_graph(code):
    c = _code_arr(code)
    L = ref_lib()
    dims = (ctypes.c_int * 2)()
    L.ref_pattern_graph(_ip(c), c.shape[0], dims, None, None, None)
    n, nnz = dims[0], dims[1]
    off = np.zeros(n + 1, np.int32)
    vl = np.zeros(n, np.int32)
    ed = np.zeros((nnz, 3), np.int32)
    L.ref_pattern_graph(_ip(c), c.shape[0], dims, _ip(off), _ip(vl), _ip(ed))
    return off, vl, ed


def ref_convert(txt_path: str, bin_path: str) -> str:
    r = subprocess.run([os.path.join(REF_DIR, "convert"), txt_path, bin_path], check=True,
                       capture_output=True, text=True)
    return r.stdout
