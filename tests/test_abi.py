"""CPU tests of the C ABI: the library loads, exports every declared symbol,
and rejects bad arguments before touching the GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

import cudaexactknn_amd as knn
from cudaexactknn_amd import _lib
from tests import pins

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
    # one-dimensional ball trees are refused before any device work
    assert L.knn_ball_build(None, 0, 1, 1, 0, ctypes.byref(h)) == _lib.KNN_ERR_UNSUPPORTED
    assert b"dims >= 2" in L.knn_last_error()
    ti = _lib.TreeInfo()
    assert L.knn_tree_get_info(None, ctypes.byref(ti)) == _lib.KNN_ERR_INVALID_ARG
    with pytest.raises(knn.KnnError):
        _lib.check(L.knn_query(None, None, 1, 1.0, None, None, 4, None))


def _sizing_failures(oracle):
    g = np.random.Generator(np.random.Philox(11))
    fails = []
    for (n, skips) in pins.SIZING_TABLE:
        if n > 65537:
            continue
        pts = g.random((n, 3), dtype=np.float32)
        for ball in (False, True):
            t = oracle.BallTree(pts, skips) if ball else oracle.KdTree(pts, skips)
            ok, msg = pins.check_sizing(t.info(), n, skips, ball)
            if not ok:
                fails.append(msg)
    return fails


def test_tree_sizing_matches_hand_written_table(oracle):
    """Oracle tree sizing (src/KdTree.cuh:56-62, src/BallTree.cuh:123-128)
    against literal values for n in {1..5, 2^16, 2^16+1} and depthSkips in
    {0, 1, 3, 40}."""
    assert _sizing_failures(oracle) == []


@pytest.mark.parametrize("mutation", [8, 9])
def test_planted_sizing_mistakes_are_caught(oracle, mutation):
    """A dropped ball clamp (8) or a log2Ceil that rounds exact powers of two
    up (9), planted in the oracle, disagrees with the table."""
    oracle.set_mutation(mutation)
    try:
        fails = _sizing_failures(oracle)
    finally:
        oracle.set_mutation(0)
    assert fails, f"mutation {mutation} passed the sizing table"


def _fake_structure(dims=3):
    """A structure object with no handle: the Python checks run before any
    call into the library, so they can be exercised without a GPU."""
    s = knn.BruteForceGpu.__new__(knn.BruteForceGpu)
    knn._Structure.__init__(s)
    s.dims, s.size, s.device = dims, 10, None
    return s


def test_python_layer_rejects_wrong_dtypes_and_sizes():
    import torch
    with pytest.raises(knn.KnnError, match="float32"):
        knn.BruteForceGpu(torch.zeros((10, 3), dtype=torch.float64))
    with pytest.raises(knn.KnnError, match="float32"):
        knn.KdTreeGpu(torch.zeros((10, 3), dtype=torch.float16))
    with pytest.raises(knn.KnnError, match="rows"):
        knn.BallTreeGpu(torch.zeros((10, 3), dtype=torch.float32), size=11)
    s = _fake_structure()
    q = torch.zeros((4, 3), dtype=torch.float32)
    ri = torch.zeros((4, 2), dtype=torch.int32)
    rd = torch.zeros((4, 2), dtype=torch.float32)
    cases = [
        (knn.Query(q.double(), 4, 1.0, ri, rd), "float32"),                    # float64 queries
        (knn.Query(q, 4, 1.0, ri.long(), rd), "int32"),                        # int64 indices
        (knn.Query(q, 4, 1.0, ri, rd.half()), "float32"),                      # half-size distances
        (knn.Query(q, 4, 1.0, ri, rd.to(torch.bfloat16)), "float32"),
        (knn.Query(q, 4, 1.0, ri, rd[:3]), "count\\*K"),
        (knn.Query(q, 4, 1.0, ri[:3], rd), "count\\*K"),                       # too few result rows
        (knn.Query(q, 5, 1.0, ri, rd), "rows"),                                # count beyond the query rows
        (knn.Query(torch.zeros((4, 2)), 4, 1.0, ri, rd), "dims"),              # wrong D
    ]
    for query, msg in cases:
        with pytest.raises(knn.KnnError, match=msg):
            s._query(query, 2, knn.QueryOpts())
    with pytest.raises(knn.KnnError, match="GPU"):                             # right types, host memory
        s._query(knn.Query(q, 4, 1.0, ri, rd), 2, knn.QueryOpts())
    it = knn.BatchItem(knn.KIND_KD, torch.zeros((8, 3)), torch.zeros((2, 3)), 4, torch.zeros((2, 4), dtype=torch.int64),
                       torch.zeros((2, 4)))
    with pytest.raises(knn.KnnError, match="int32"):
        knn.Batch([it])


def test_tuned_table_is_compiled_into_the_library(tmp_path):
    """The library was built from the committed autotuner table, and
    knn_tuned_lookup returns every committed entry (a host-only call)."""
    from cudaexactknn_amd import autotune
    from cudaexactknn_amd import build as B
    inc = tmp_path / "table.inc"
    n = B.gen_tuned_table(dst=str(inc))
    assert inc.read_text() == open(B.TUNED_INC).read(), "rebuild: the library's table is stale"
    table = autotune.load_table()
    assert n == len(table)
    for key, e in table.items():
        cls, d, k, nb = key.split(":")
        got = autotune.library_lookup(cls, int(d[1:]), int(k[1:]), 1 << int(nb[3:]))
        assert got is not None, key
        assert got["BLOCK_SIZE"] == int(e["best"].get("BLOCK_SIZE", 0) or 0), key
        assert got["leaf_span_delta"] == int(e["best"].get("leaf_span_delta", 0) or 0), key
        assert got["tile"] == int(e["best"].get("tile", 0) or 0), key
    # more than one size bucket from every tuned size: no entry
    assert all(int(key.split(":")[3][3:]) < 29 for key in table)
    assert autotune.library_lookup("kd", 3, 16, 1 << 30) is None
