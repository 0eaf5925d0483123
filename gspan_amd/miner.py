"""Python front end of the MI355X gSpan path.

``fsm(g, k, min_support)`` mirrors the reference's ``void fsm(const CSRGraph&, int k,
int min_support)`` (gspan.h:41, gspan.cu:1171-1199) but returns the reported patterns
instead of printing them; ``format_report`` gives the exact lines the reference prints
(gspan.cu:67-77).  Everything runs in libgspan_amd.so on the GPU; this module only
marshals arguments.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from .graph import CSRGraph

__all__ = ["Miner", "fsm", "format_report"]


def _ip(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _fp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _code_arr(code) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(code, dtype=np.int32).reshape(-1, 5))


def _layers(layers):
    arrs = [np.ascontiguousarray(np.asarray(x, dtype=np.int32).reshape(-1, 2)) for x in layers]
    ptrs = (ctypes.POINTER(ctypes.c_int) * max(len(arrs), 1))(*[_ip(a) for a in arrs])
    counts = (ctypes.c_longlong * max(len(arrs), 1))(*[a.shape[0] for a in arrs])
    return arrs, ptrs, counts


class Miner:
    """One device context holding one data graph (``gs_create`` .. ``gs_destroy``)."""

    def __init__(self, g: CSRGraph, mode: str = "batched", mni: str = "gpu", device: int = 0,
                 timing: bool = False, budget_bytes: int = 0, chunk_rows: int = 0,
                 two_pass: bool = False, sets: str = "auto", autotune: bool = True,
                 tile: int = 0, groups: str = "auto", two_edge: str = "auto"):
        self._lib = L.lib()
        opts = L.GsOpts()
        self._lib.gs_default_opts(ctypes.byref(opts))
        opts.mode = {"plain": L.GS_MODE_PLAIN, "batched": L.GS_MODE_BATCHED,
                     "predicted": L.GS_MODE_PREDICTED}[mode]
        opts.mni_mode = {"gpu": L.GS_MNI_GPU, "omp-literal": L.GS_MNI_OMP_LITERAL}[mni]
        opts.device = device
        opts.timing = 1 if timing else 0
        opts.budget_bytes = budget_bytes
        opts.chunk_rows = chunk_rows
        opts.two_pass = 1 if two_pass else 0
        opts.set_kind = {"auto": L.GS_SETS_AUTO, "bits": L.GS_SETS_BITS,
                         "bytes": L.GS_SETS_BYTES}[sets]
        opts.autotune = 1 if autotune else 0
        opts.tile = tile
        opts.groups = {"auto": L.GS_GROUPS_AUTO, "on": L.GS_GROUPS_ON,
                       "off": L.GS_GROUPS_OFF}[groups]
        opts.two_edge = {"auto": L.GS_TWO_EDGE_AUTO, "off": L.GS_TWO_EDGE_OFF}[two_edge]
        self.mni_mode = opts.mni_mode
        self.graph = g
        h = ctypes.c_void_p()
        rc = self._lib.gs_create(_ip(g.offsets), _ip(g.vlabels), _ip(g.edges), g.nr_nodes,
                                 g.nr_edges, ctypes.byref(opts), ctypes.byref(h))
        if rc != L.GS_OK:
            raise L.GsError(rc, "gs_create failed (see stderr)")
        self._h = h

    def _check(self, rc: int):
        if rc != L.GS_OK:
            msg = self._lib.gs_last_error(self._h)
            raise L.GsError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.gs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def mine_raw(self, k: int, min_support: int):
        """Runs gs_mine and returns its results as flat arrays in the reference's report
        order: (codes int32[E, 5], lens int32[P], supports int32[P])."""
        self._check(self._lib.gs_mine(self._h, k, min_support, L.REPORT_CB(0), None))
        return self.results()

    def results(self):
        n, e = ctypes.c_longlong(), ctypes.c_longlong()
        self._check(self._lib.gs_result_count(self._h, ctypes.byref(n), ctypes.byref(e)))
        codes = np.empty((e.value, 5), np.int32)
        lens = np.empty(n.value, np.int32)
        sups = np.empty(n.value, np.int32)
        self._check(self._lib.gs_result_copy(self._h, _ip(codes), _ip(lens), _ip(sups)))
        return codes, lens, sups

    @staticmethod
    def to_patterns(codes, lens, sups):
        """Flat result arrays -> [(code tuple of 5-tuples, support)]."""
        out = []
        at = 0
        rows = codes.tolist()
        for m, sup in zip(lens.tolist(), sups.tolist()):
            out.append((tuple(tuple(r) for r in rows[at:at + m]), sup))
            at += m
        return out

    def mine(self, k: int, min_support: int):
        """All frequent connected patterns with at most k edges, in the reference's
        report order: [(code tuple of 5-tuples, support)]."""
        return self.to_patterns(*self.mine_raw(k, min_support))

    def support(self, code, layers, mni: int | None = None) -> int:
        c = _code_arr(code)
        arrs, ptrs, counts = _layers(layers)
        out = ctypes.c_int()
        self._check(self._lib.gs_support(self._h, _ip(c), c.shape[0], ptrs, counts,
                                         self.mni_mode if mni is None else mni,
                                         ctypes.byref(out)))
        return out.value

    def extend(self, code, layers, mni: int | None = None):
        """One extension step: [(edge5, rows int32[count, 2], canonical, support)] in the
        reference's child order."""
        c = _code_arr(code)
        arrs, ptrs, counts = _layers(layers)
        out = []

        def cb(ep, rp, cnt, canon, sup, _u):
            e5 = tuple(ep[j] for j in range(5))
            rows = (np.ctypeslib.as_array(rp, shape=(cnt * 2,)).reshape(cnt, 2).copy()
                    if cnt else np.zeros((0, 2), np.int32))
            out.append((e5, rows, bool(canon), sup))

        total = ctypes.c_longlong()
        self._check(self._lib.gs_extend(self._h, _ip(c), c.shape[0], ptrs, counts,
                                        self.mni_mode if mni is None else mni, L.CHILD_CB(cb),
                                        None, ctypes.byref(total)))
        return out

    def is_canonical(self, codes) -> list[bool]:
        codes = [_code_arr(c) for c in codes]
        if not codes:
            return []
        flat = np.ascontiguousarray(np.concatenate(codes, axis=0))
        lens = np.array([c.shape[0] for c in codes], dtype=np.int32)
        flags = np.zeros(len(codes), dtype=np.int32)
        self._check(self._lib.gs_is_canonical_batch(self._h, _ip(flat), _ip(lens), len(codes),
                                                    _ip(flags)))
        return [bool(x) for x in flags]

    def prune_graph(self, min_sup: int) -> CSRGraph:
        """prune_graph (graph.cc:216-266) of this context's graph, computed on the GPU."""
        g = self.graph
        off = np.zeros(g.nr_nodes + 1, np.int32)
        vl = np.zeros(max(g.nr_nodes, 1), np.int32)
        ed = np.zeros((max(g.nr_edges, 1), 3), np.int32)
        n, nnz = ctypes.c_int(), ctypes.c_int()
        self._check(self._lib.gs_prune_graph(self._h, min_sup, ctypes.byref(n), ctypes.byref(nnz),
                                             _ip(off), _ip(vl), _ip(ed)))
        return CSRGraph(off[:n.value + 1].copy(), vl[:n.value].copy(), ed[:nnz.value].copy())

    # ---- child-count predictor
    def predictor_log(self, on: bool = True, capacity: int = 1 << 20):
        """Log (features, log1p(count)) of every candidate child key of exact count passes."""
        self._check(self._lib.gs_predictor_log(self._h, 1 if on else 0, capacity))

    def predictor_samples(self) -> int:
        n = ctypes.c_longlong()
        self._check(self._lib.gs_predictor_samples(self._h, ctypes.byref(n)))
        return n.value

    def predictor_train(self, epochs: int = 300, lr: float = 3e-3) -> float:
        loss = ctypes.c_float()
        self._check(self._lib.gs_predictor_train(self._h, epochs, lr, ctypes.byref(loss)))
        return loss.value

    def predictor_weights(self) -> np.ndarray:
        w = np.zeros(L.GS_MLP_PARAMS, np.float32)
        self._check(self._lib.gs_predictor_get_weights(self._h, _fp(w)))
        return w

    def set_predictor_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.shape == (L.GS_MLP_PARAMS,)
        self._check(self._lib.gs_predictor_set_weights(self._h, _fp(w)))

    def mlp_forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, L.GS_MLP_FEATURES)
        y = np.zeros(x.shape[0], np.float32)
        self._check(self._lib.gs_mlp_forward(self._h, _fp(x), x.shape[0], _fp(y)))
        return y

    def mlp_gradients(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, L.GS_MLP_FEATURES)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(L.GS_MLP_PARAMS, np.float32)
        loss = ctypes.c_float()
        self._check(self._lib.gs_mlp_gradients(self._h, _fp(x), _fp(t), x.shape[0], _fp(g),
                                               ctypes.byref(loss)))
        return g, loss.value

    def tuner_report(self) -> dict:
        """The launch-shape tuner's per-bucket choice and time per item of each shape."""
        import json
        n = ctypes.c_int()
        self._check(self._lib.gs_tuner_report(self._h, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value + 1)
        self._check(self._lib.gs_tuner_report(self._h, buf, n.value + 1, ctypes.byref(n)))
        return json.loads(buf.value.decode())

    def stats(self) -> dict:
        s = L.GsStats()
        self._check(self._lib.gs_get_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    def reset_stats(self):
        self._check(self._lib.gs_reset_stats(self._h))


def fsm(g: CSRGraph, k: int, min_support: int, mode: str = "batched", mni: str = "gpu"):
    """fsm(g, k, min_support) (gspan.cu:1171): the reported patterns, in order."""
    with Miner(g, mode=mode, mni=mni) as m:
        return m.mine(k, min_support)


def format_report(patterns) -> str:
    """The reference's stdout lines (gspan.cu:67-77): ``(from to fl el tl)...: support``."""
    lines = ["".join("(%d %d %d %d %d)" % e for e in code) + ": %d" % sup
             for code, sup in patterns]
    return "\n".join(lines) + ("\n" if lines else "")
