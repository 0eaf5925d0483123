"""gs_is_canonical_batch_noalloc (no allocation in the call: its device staging lives in a
static device array) against gs_is_canonical_batch, called through the C-ABI.

Two calls in a row, the second with fewer codes than the first (the shared array keeps the
first call's contents), then a call past GS_NOALLOC_MAX_CODES, which must be rejected.
"""
import ctypes

import numpy as np
import pytest

import brute
from gspan_amd import _lib as L
from gspan_amd.graph import synthetic
from gspan_amd.miner import Miner

pytestmark = pytest.mark.gpu


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _codes(rng, count):
    """Random DFS codes of 1..7 edges, with labels far apart (ranks must not matter), plus
    the single-label 7-star rooted at a leaf (5040 projections) and at its centre."""
    out = []
    while len(out) < count - 2:
        c = brute.random_code(rng, int(rng.integers(1, 8)), int(rng.integers(1, 4)),
                              int(rng.integers(1, 3)))
        if c[0][2] <= c[0][4]:
            # spread the labels: 0, 1, 2 -> 5, 105, 205 (vertex), 0, 1 -> 7, 1007 (edge)
            out.append(tuple((f, t, -1 if fl < 0 else 5 + 100 * fl, 7 + 1000 * el,
                              -1 if tl < 0 else 5 + 100 * tl) for f, t, fl, el, tl in c))
    leaf = ((0, 1, 4, 9, 4),) + tuple((1, j, -1, 9, 4) for j in range(2, 8))
    centre = ((0, 1, 4, 9, 4),) + tuple((0, j, -1, 9, 4) for j in range(2, 8))
    return out + [leaf, centre]


def _flat(codes):
    flat = np.ascontiguousarray(np.array([e for c in codes for e in c], np.int32).reshape(-1, 5))
    lens = np.array([len(c) for c in codes], np.int32)
    return flat, lens


def _noalloc(lib, h, codes):
    flat, lens = _flat(codes)
    flags = np.full(len(codes), -7, np.int32)
    rc = lib.gs_is_canonical_batch_noalloc(h, _ip(flat), _ip(lens), len(codes), _ip(flags))
    return rc, flags


def test_noalloc_matches_canonical_batch():
    lib = L.lib()
    rng = np.random.default_rng(11)
    g = synthetic("er", n=50, m=100, n_vlabels=2, seed=1)
    first = _codes(rng, 3000)
    second = _codes(rng, 700)
    with Miner(g) as mn:
        h = mn._h
        for name, codes in (("first", first), ("second, smaller", second)):
            print(f"[gpu-case] noalloc canonicality, {name}: {len(codes)} codes", flush=True)
            want = np.array(mn.is_canonical(codes), np.int32)
            rc, got = _noalloc(lib, h, codes)
            assert rc == L.GS_OK, lib.gs_last_error(h)
            np.testing.assert_array_equal(got, want)
            assert got[-2:].tolist() == [1, 0]  # leaf-rooted 7-star canonical, centre not
            assert 0 < got.sum() < len(codes)
        print("[gpu-case] noalloc canonicality past the limit", flush=True)
        too_many = [first[0]] * (L.GS_NOALLOC_MAX_CODES + 1)
        rc, flags = _noalloc(lib, h, too_many)
        assert rc == L.GS_ERR_LIMIT
        assert b"GS_NOALLOC_MAX_CODES" in lib.gs_last_error(h)
        assert (flags == -7).all()  # nothing written
        rc, got = _noalloc(lib, h, second)  # still usable afterwards
        assert rc == L.GS_OK
        np.testing.assert_array_equal(got, np.array(mn.is_canonical(second), np.int32))
