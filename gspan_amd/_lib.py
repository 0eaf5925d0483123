"""ctypes binding of libgspan_amd.so (C-ABI declared in include/gspan_amd.h).

The library is built in-tree by ``gspan_amd.build`` (``__graft_entry__.build()``).  There
is no fallback: if the shared library is missing or fails to load, every call raises.
"""
from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "lib", "libgspan_amd.so")
HEADER = os.path.join(ROOT, "include", "gspan_amd.h")

GS_OK = 0
GS_ERR_INVALID = 1
GS_ERR_HIP = 2
GS_ERR_BUDGET = 3
GS_ERR_LIMIT = 4
GS_ERR_CALLBACK = 5
GS_ERR_INTERNAL = 6

GS_MODE_PLAIN = 0
GS_MODE_BATCHED = 1
GS_MODE_PREDICTED = 2
GS_MLP_FEATURES = 32
GS_MLP_PARAMS = 20865
GS_MNI_GPU = 0
GS_MNI_OMP_LITERAL = 1
GS_SETS_AUTO = 0
GS_SETS_BITS = 1
GS_SETS_BYTES = 2
GS_GROUPS_AUTO = 0
GS_GROUPS_ON = 1
GS_GROUPS_OFF = 2
GS_TWO_EDGE_AUTO = 0
GS_TWO_EDGE_OFF = 1
GS_MAX_PATTERN_EDGES = 7
GS_NOALLOC_MAX_CODES = 8192


class GsOpts(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_int),
        ("mni_mode", ctypes.c_int),
        ("device", ctypes.c_int),
        ("timing", ctypes.c_int),
        ("budget_bytes", ctypes.c_ulonglong),
        ("chunk_rows", ctypes.c_longlong),
        ("two_pass", ctypes.c_int),
        ("set_kind", ctypes.c_int),
        ("autotune", ctypes.c_int),
        ("tile", ctypes.c_int),
        ("groups", ctypes.c_int),
        ("two_edge", ctypes.c_int),
    ]


class GsStats(ctypes.Structure):
    _fields_ = [
        ("extended", ctypes.c_longlong),
        ("candidates", ctypes.c_longlong),
        ("canonical", ctypes.c_longlong),
        ("reported", ctypes.c_longlong),
        ("batches", ctypes.c_longlong),
        ("items", ctypes.c_longlong),
        ("parent_rows", ctypes.c_longlong),
        ("fill_launches", ctypes.c_longlong),
        ("count_launches", ctypes.c_longlong),
        ("alg_bytes_fill", ctypes.c_double),
        ("alg_bytes_count", ctypes.c_double),
        ("t_fill_ms", ctypes.c_double),
        ("t_count_ms", ctypes.c_double),
        ("t_mine_ms", ctypes.c_double),
        ("peak_device_bytes", ctypes.c_ulonglong),
        ("predicted_steps", ctypes.c_longlong),
        ("fallbacks", ctypes.c_longlong),
        ("spilled_rows", ctypes.c_longlong),
        ("t_sync_ms", ctypes.c_double),
        ("t_order_ms", ctypes.c_double),
        ("t_plan_ms", ctypes.c_double),
        ("two_edge_levels", ctypes.c_longlong),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


REPORT_CB = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                             ctypes.c_int, ctypes.c_void_p)
CHILD_CB = ctypes.CFUNCTYPE(None, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                            ctypes.c_longlong, ctypes.c_int, ctypes.c_int, ctypes.c_void_p)

_lib = None


class GsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[gspan_amd error {code}] {msg}")
        self.code = code


def declared_symbols() -> list[str]:
    """Names of the functions declared in include/gspan_amd.h."""
    with open(HEADER) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(gs_\w+)\s*\(", text, re.M)))


def lib():
    """Loads libgspan_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    c_int, c_ll, vp = ctypes.c_int, ctypes.c_longlong, ctypes.c_void_p
    L.gs_version.restype = ctypes.c_char_p
    L.gs_version.argtypes = []
    L.gs_default_opts.argtypes = [P(GsOpts)]
    L.gs_create.argtypes = [P(c_int), P(c_int), P(c_int), c_int, c_int, P(GsOpts), P(vp)]
    L.gs_mine.argtypes = [vp, c_int, c_int, REPORT_CB, vp]
    L.gs_support.argtypes = [vp, P(c_int), c_int, P(P(c_int)), P(c_ll), c_int, P(c_int)]
    L.gs_extend.argtypes = [vp, P(c_int), c_int, P(P(c_int)), P(c_ll), c_int, CHILD_CB, vp,
                            P(c_ll)]
    L.gs_is_canonical_batch.argtypes = [vp, P(c_int), P(c_int), c_int, P(c_int)]
    L.gs_is_canonical_batch_noalloc.argtypes = [vp, P(c_int), P(c_int), c_int, P(c_int)]
    c_f = ctypes.c_float
    L.gs_predictor_log.argtypes = [vp, c_int, c_ll]
    L.gs_predictor_samples.argtypes = [vp, P(c_ll)]
    L.gs_predictor_train.argtypes = [vp, c_int, c_f, P(c_f)]
    L.gs_predictor_get_weights.argtypes = [vp, P(c_f)]
    L.gs_predictor_set_weights.argtypes = [vp, P(c_f)]
    L.gs_mlp_forward.argtypes = [vp, P(c_f), c_ll, P(c_f)]
    L.gs_mlp_gradients.argtypes = [vp, P(c_f), P(c_f), c_ll, P(c_f), P(c_f)]
    L.gs_prune_graph.argtypes = [vp, c_int, P(c_int), P(c_int), P(c_int), P(c_int), P(c_int)]
    L.gs_tuner_report.argtypes = [vp, ctypes.c_char_p, c_int, P(c_int)]
    L.gs_text_parse.argtypes = [ctypes.c_char_p, c_ll, c_int, P(vp), P(c_int), P(c_int), P(c_ll)]
    L.gs_text_copy.argtypes = [vp, P(c_int), P(c_int), P(c_int)]
    L.gs_text_free.argtypes = [vp]
    L.gs_text_free.restype = None
    L.gs_result_count.argtypes = [vp, P(c_ll), P(c_ll)]
    L.gs_result_copy.argtypes = [vp, P(c_int), P(c_int), P(c_int)]
    L.gs_get_stats.argtypes = [vp, P(GsStats)]
    L.gs_reset_stats.argtypes = [vp]
    L.gs_last_error.argtypes = [vp]
    L.gs_last_error.restype = ctypes.c_char_p
    L.gs_destroy.argtypes = [vp]
    for name in ("gs_default_opts", "gs_create", "gs_mine", "gs_support", "gs_extend",
                 "gs_is_canonical_batch", "gs_is_canonical_batch_noalloc", "gs_result_count", "gs_result_copy", "gs_get_stats",
                 "gs_reset_stats", "gs_destroy", "gs_predictor_log", "gs_predictor_samples",
                 "gs_predictor_train", "gs_predictor_get_weights", "gs_predictor_set_weights",
                 "gs_mlp_forward", "gs_mlp_gradients", "gs_prune_graph", "gs_tuner_report",
                 "gs_text_parse", "gs_text_copy"):
        getattr(L, name).restype = c_int
    _lib = L
    return L
