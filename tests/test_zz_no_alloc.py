"""knn_query_noalloc through the C ABI against knn_query on the same handle.

Each structure answers two calls in a row, the second with fewer queries
(the static scratch keeps the first call's contents), and every row of both
calls must equal knn_query's bit for bit.  Covers the tree ordering scratch
(kd and ball, wave-cooperative and unsorted per-lane), the brute-force split
scan (few queries over many points: partial keys in the scratch), heap and
array rows, a finite radius, and the size limit's error.
"""
import ctypes

import numpy as np
import pytest

from cudaexactknn_amd import _lib

pytestmark = pytest.mark.gpu


def _handle(kind, p, n, d, skips):
    L = _lib.load()
    h = ctypes.c_void_p(0)
    if kind == "bf":
        _lib.check(L.knn_bf_create(p.data_ptr(), n, d, ctypes.byref(h)))
    elif kind == "kd":
        _lib.check(L.knn_kd_build(p.data_ptr(), n, d, d, skips, 0, ctypes.byref(h)))
    else:
        _lib.check(L.knn_ball_build(p.data_ptr(), n, d, skips, 0, ctypes.byref(h)))
    return h


def _call(fn, h, q, k, md, opts):
    import torch
    ri = torch.full((q.shape[0], k), 7, dtype=torch.int32, device=q.device)
    rd = torch.full((q.shape[0], k), 7.0, dtype=torch.float32, device=q.device)
    _lib.check(fn(h, q.data_ptr(), q.shape[0], float(md), ri.data_ptr(), rd.data_ptr(), k, ctypes.byref(opts)))
    torch.cuda.synchronize()
    return ri.cpu().numpy(), rd.cpu().numpy().view(np.uint32)


CASES = [
    # kind, points, dims, skips, queries (first, second), K, heap, flags
    ("kd", 200000, 3, 0, (50000, 3001), 16, 0, 0),
    ("kd", 200000, 3, 0, (40000, 777), 8, 1, _lib.KNN_FLAG_NO_QUERY_SORT),
    ("ball", 150000, 4, 1, (30000, 2049), 16, 0, 0),
    ("kd", 100000, 8, 2, (20000, 513), 32, 0, 0),
    ("bf", 120000, 3, 0, (600, 37), 16, 0, 0),
    ("bf", 60000, 32, 0, (300, 5), 32, 1, 0),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-d{c[2]}-k{c[5]}-h{c[6]}-f{c[7]}" for c in CASES])
def test_noalloc_equals_knn_query_on_two_calls(cuda, case):
    import torch
    kind, n, d, skips, (q1, q2), k, heap, flags = case
    g = np.random.Generator(np.random.Philox(n + d + k))
    pts = g.random((n, d), dtype=np.float32)
    pts[5] = pts[6]
    p = torch.from_numpy(pts).to(cuda)
    h = _handle(kind, p, n, d, skips)
    L = _lib.load()
    try:
        for nq, md in ((q1, np.inf), (q2, float(np.float32(0.02 * d)))):
            qs = g.random((nq, d), dtype=np.float32)
            qs[0] = pts[6]
            q = torch.from_numpy(qs).to(cuda)
            opts = _lib.QueryOpts()
            opts.heap, opts.flags = heap, flags
            want = _call(L.knn_query, h, q, k, md, opts)
            got = _call(L.knn_query_noalloc, h, q, k, md, opts)
            assert np.array_equal(got[0], want[0]), f"{case} nq={nq}: indices differ"
            assert np.array_equal(got[1], want[1]), f"{case} nq={nq}: distance bits differ"
            assert (want[0] >= 0).any()
    finally:
        L.knn_destroy(h)


def test_noalloc_rejects_batches_above_its_scratch(cuda):
    """6.5M queries need more ordering scratch than the static maximum: the
    call fails with a clear error before any launch (the buffers are real and
    large enough, so nothing could run out of bounds either way)."""
    import torch
    pts = np.random.Generator(np.random.Philox(1)).random((1000, 3), dtype=np.float32)
    p = torch.from_numpy(pts).to(cuda)
    h = _handle("kd", p, 1000, 3, 0)
    L = _lib.load()
    try:
        nq = 6_500_000
        q = torch.zeros((nq, 3), dtype=torch.float32, device=cuda)
        ri = torch.empty((nq, 1), dtype=torch.int32, device=cuda)
        rd = torch.empty((nq, 1), dtype=torch.float32, device=cuda)
        st = L.knn_query_noalloc(h, q.data_ptr(), nq, float("inf"), ri.data_ptr(), rd.data_ptr(), 1, None)
        assert st == _lib.KNN_ERR_INVALID_ARG
        assert b"static maximum" in L.knn_last_error()
        opts = _lib.QueryOpts()
        opts.form = _lib.KNN_FORM_LEARNED
        st = L.knn_query_noalloc(h, q.data_ptr(), 10, float("inf"), ri.data_ptr(), rd.data_ptr(), 1,
                                 ctypes.byref(opts))
        assert st == _lib.KNN_ERR_UNSUPPORTED
    finally:
        L.knn_destroy(h)
