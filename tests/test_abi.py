"""CPU tests of the C ABI: the library loads, exports every declared symbol,
and rejects bad arguments before touching the GPU."""
import ctypes
import os
import re

import pytest

import cudaexactknn_amd as knn
from cudaexactknn_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    src = open(os.path.join(ROOT, "include", "knn_mi355x.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(knn_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_exported_list():
    assert _declared_symbols() == sorted(_lib.EXPORTED)


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    for name in _declared_symbols():
        assert hasattr(L, name), name


def test_version_and_struct_sizes():
    assert knn.version() >= 0x000100
    assert ctypes.sizeof(_lib.QueryOpts) == 8 * 4 + 8
    assert ctypes.sizeof(_lib.TreeInfo) == 12 * 4


def test_null_handle_and_bad_args_fail_cleanly():
    L = _lib.load()
    assert L.knn_destroy(None) == _lib.KNN_OK
    h = ctypes.c_void_p(0)
    assert L.knn_bf_create(None, 10, 3, ctypes.byref(h)) == _lib.KNN_ERR_INVALID_ARG
    assert b"NULL" in L.knn_last_error()
    assert L.knn_kd_build(None, 0, 0, 0, 0, 0, ctypes.byref(h)) == _lib.KNN_ERR_INVALID_ARG
    assert L.knn_kd_build(None, 0, 3, 4, 0, 0, ctypes.byref(h)) == _lib.KNN_ERR_INVALID_ARG
    assert L.knn_ball_build(None, 5, 3, 1, 2048, ctypes.byref(h)) == _lib.KNN_ERR_INVALID_ARG
    ti = _lib.TreeInfo()
    assert L.knn_tree_get_info(None, ctypes.byref(ti)) == _lib.KNN_ERR_INVALID_ARG
    with pytest.raises(knn.KnnError):
        _lib.check(L.knn_query(None, None, 1, 1.0, None, None, 4, None))


def test_tree_sizing_matches_reference_formula(oracle):
    """tree sizing (src/KdTree.cuh:56-62, src/BallTree.cuh:123-128) as exported
    by the oracle for edge sizes; the library uses the same rule."""
    for n in (1, 2, 3, 4, 5, 1000, 1 << 16, (1 << 16) + 1):
        for skips in (0, 1, 3, 40):
            L = oracle.log2_ceil(n)
            s_kd = min(skips, L - 1)
            s_ball = min(max(skips, 1), L - 1)
            for s in (s_kd, s_ball):
                assert (1 << (L - s)) - 1 >= 1
                assert (1 << L) >= n
