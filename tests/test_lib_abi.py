"""The C-ABI library loads and exports every function include/gspan_amd.h declares.
Only calls that fail argument validation before touching the GPU run here."""
import ctypes

import numpy as np
import pytest

from gspan_amd import _lib as L


def test_library_exports_declared_symbols():
    lib = L.lib()
    names = L.declared_symbols()
    assert "gs_mine" in names and "gs_create" in names and len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
    assert b"gfx950" in lib.gs_version()


def test_default_opts():
    o = L.GsOpts()
    assert L.lib().gs_default_opts(ctypes.byref(o)) == L.GS_OK
    assert o.mode == L.GS_MODE_BATCHED and o.mni_mode == L.GS_MNI_GPU
    assert o.autotune == 1 and o.tile == 0 and o.set_kind == L.GS_SETS_AUTO
    n = ctypes.c_int()
    assert L.lib().gs_tuner_report(None, None, 0, ctypes.byref(n)) == L.GS_ERR_INVALID


def _create(offsets, vlabels, edges, n, nnz):
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    h = ctypes.c_void_p()
    rc = L.lib().gs_create(ip(offsets), ip(vlabels), ip(edges), n, nnz, None, ctypes.byref(h))
    return rc, h


def test_create_rejects_unsorted_adjacency():
    off = np.array([0, 2, 3, 4], np.int32)
    vl = np.zeros(3, np.int32)
    ed = np.array([[0, 2, 0], [0, 1, 0], [1, 0, 0], [2, 0, 0]], np.int32)
    rc, h = _create(off, vl, ed, 3, 4)
    assert rc == L.GS_ERR_INVALID and not h.value


def test_create_rejects_bad_offsets_and_from():
    vl = np.zeros(2, np.int32)
    ed = np.array([[0, 1, 0], [1, 0, 0]], np.int32)
    assert _create(np.array([0, 1, 3], np.int32), vl, ed, 2, 2)[0] == L.GS_ERR_INVALID
    bad = np.array([[1, 1, 0], [1, 0, 0]], np.int32)
    assert _create(np.array([0, 1, 2], np.int32), vl, bad, 2, 2)[0] == L.GS_ERR_INVALID
    assert _create(np.array([0, 1, 2], np.int32), vl, ed, -1, 2)[0] == L.GS_ERR_INVALID


def test_null_context_calls_fail_cleanly():
    lib = L.lib()
    assert lib.gs_mine(None, 3, 1, L.REPORT_CB(lambda *a: 0), None) == L.GS_ERR_INVALID
    assert lib.gs_destroy(None) == L.GS_OK
    assert lib.gs_last_error(None) == b"null context"


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(L, "_lib", None)
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(RuntimeError, match="missing"):
        L.lib()
