"""knn_query enqueued on a stream and captured into a HIP graph.

The reference runs batchKnn on the default stream only (its one stream hook
is the Query it hands the kernel, src/KnnTrees.cuh:209-243).  Here a query on
a caller's stream allocates nothing once knn_query_reserve has sized the
handle, never waits for the device, and so can be captured.  Each case:
reserve, capture one call (torch.cuda.graph, global capture mode, so any
allocation or synchronisation inside the call fails the capture), then replay
the graph three times on new query values copied into the captured input,
with an eager synchronous knn_query on the same handle between replays; every
replay and every eager call must equal the oracle's brute force bit for bit.
A second graph of fewer queries than reserved checks that the reservation
covers every smaller batch (the brute-force forms split the points further
when fewer queries fill the GPU).

Covers kd and ball trees (wave-cooperative, sorted queries), the plain brute
force with its split scan, the bf16 form (seeded screen, and an all-equidistant
set whose candidate lists overflow and are re-answered on the device), the
learned radius's in-kernel second pass, and knn_query_noalloc.
"""
import ctypes

import numpy as np
import pytest

import cudaexactknn_amd as knn
from cudaexactknn_amd import _lib

pytestmark = pytest.mark.gpu


def _rows(ri, rd):
    return ri.cpu().numpy(), rd.cpu().numpy().view(np.uint32)


def _same(got, want, what):
    assert np.array_equal(got[0], want[0]), f"{what}: indices differ"
    assert np.array_equal(got[1], want[1].view(np.uint32)), f"{what}: distance bits differ"


def _case_points(kind, seed):
    g = np.random.Generator(np.random.Philox(seed))
    if kind == "overflow":
        v = g.standard_normal((6000, 32))
        return (v / np.linalg.norm(v, axis=1, keepdims=True) * 4.0).astype(np.float32), g
    n, d = {"kd": (60000, 3), "ball": (40000, 4), "bf": (80000, 3), "bf16": (65536, 32),
            "learned": (60000, 3), "learned_radius": (60000, 3), "noalloc": (60000, 3)}[kind]
    pts = (g.standard_normal((n, d)) * np.exp(g.uniform(-0.5, 0.5, d))).astype(np.float32)
    pts[7] = pts[8]
    return pts, g


def _queries(kind, pts, g, nq):
    d = pts.shape[1]
    if kind == "overflow":
        qs = np.zeros((nq, d), np.float32)
        qs[1::2] = pts[g.integers(0, len(pts), len(qs[1::2]))]
        return qs
    qs = (g.standard_normal((nq, d)) * 1.1).astype(np.float32)
    qs[0] = pts[8]
    return qs


CASES = [
    # name, structure, form, k, heap, queries reserved (and captured), smaller capture
    ("kd", "kd", _lib.KNN_FORM_PLAIN, 16, 0, 4096, 1000),
    ("ball", "ball", _lib.KNN_FORM_PLAIN, 8, 1, 3000, 640),
    ("bf", "bf", _lib.KNN_FORM_PLAIN, 16, 0, 600, 70),
    ("bf16", "bf", _lib.KNN_FORM_BF16, 32, 0, 512, 100),
    ("overflow", "bf", _lib.KNN_FORM_BF16, 32, 1, 64, 9),
    ("learned", "kd", _lib.KNN_FORM_LEARNED, 16, 0, 4096, 1000),
    ("learned_radius", "kd", _lib.KNN_FORM_LEARNED, 16, 0, 4096, 1000),
    ("noalloc", "kd", _lib.KNN_FORM_PLAIN, 16, 0, 4096, 1000),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_captured_query_replays_match_oracle(oracle, cuda, case):
    import torch
    name, kind, form, k, heap, nq, nq_small = case
    pts, g = _case_points(name, 31 + len(name))
    d = pts.shape[1]
    p = torch.from_numpy(pts).to(cuda)
    if kind == "kd":
        t = knn.KdTreeGpu(p, depth_skips=0)
    elif kind == "ball":
        t = knn.BallTreeGpu(p, depth_skips=1)
    else:
        t = knn.BruteForceGpu(p)
    L = _lib.load()
    fn = L.knn_query_noalloc if name == "noalloc" else L.knn_query
    try:
        if name.startswith("learned"):
            t.train_radius(torch.from_numpy(_queries(name, pts, g, 20000)).to(cuda), k)
        flags = _lib.KNN_FLAG_LEARNED_RADIUS if name == "learned_radius" else 0
        stream = torch.cuda.Stream(device=cuda)
        opts = _lib.QueryOpts()
        opts.form, opts.heap, opts.flags = form, heap, flags
        opts.stream = stream.cuda_stream
        sync = _lib.QueryOpts()
        sync.form, sync.heap, sync.flags = form, heap, flags
        _lib.check(L.knn_query_reserve(t._h, nq, k, ctypes.byref(opts)))
        for m in (nq, nq_small):
            print(f"[capture] {name} nq={m} k={k} heap={heap}", flush=True)
            qbuf = torch.zeros((m, d), dtype=torch.float32, device=cuda)
            ri = torch.empty((m, k), dtype=torch.int32, device=cuda)
            rd = torch.empty((m, k), dtype=torch.float32, device=cuda)
            if name == "noalloc":
                # the first call on a device looks up the static scratch's address
                _lib.check(fn(t._h, qbuf.data_ptr(), m, float("inf"), ri.data_ptr(), rd.data_ptr(), k,
                              ctypes.byref(sync)))
            torch.cuda.synchronize()
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph, stream=stream):
                st = fn(t._h, qbuf.data_ptr(), m, float("inf"), ri.data_ptr(), rd.data_ptr(), k, ctypes.byref(opts))
            _lib.check(st)
            ei = torch.empty((m, k), dtype=torch.int32, device=cuda)
            ed = torch.empty((m, k), dtype=torch.float32, device=cuda)
            for r in range(3):
                qs = _queries(name, pts, g, m)
                qbuf.copy_(torch.from_numpy(qs))
                ri.fill_(7)
                rd.fill_(7.0)
                graph.replay()
                torch.cuda.synchronize()
                want = oracle.bf_knn(pts, qs, k, heap=bool(heap))
                _same(_rows(ri, rd), want, f"{name} nq={m} replay {r}")
                # eager work on the same handle between replays
                other = _queries(name, pts, g, m)
                eq = torch.from_numpy(other).to(cuda)
                _lib.check(L.knn_query(t._h, eq.data_ptr(), m, float("inf"), ei.data_ptr(), ed.data_ptr(), k,
                                       ctypes.byref(sync)))
                _same(_rows(ei, ed), oracle.bf_knn(pts, other, k, heap=bool(heap)), f"{name} nq={m} eager {r}")
            if name == "overflow":
                assert t.last_query_stats()["rerun_queries"] >= 1
            del graph
    finally:
        t.close()


def test_capture_without_reserve_fails_cleanly(cuda):
    """A captured query that would have to grow the handle's scratch returns
    KNN_ERR_INVALID_ARG naming knn_query_reserve, before any allocation or
    launch; the handle still answers eager queries afterwards."""
    import torch
    g = np.random.Generator(np.random.Philox(3))
    pts = g.random((5000, 3), dtype=np.float32)
    t = knn.KdTreeGpu(torch.from_numpy(pts).to(cuda))
    L = _lib.load()
    try:
        stream = torch.cuda.Stream(device=cuda)
        opts = _lib.QueryOpts()
        opts.stream = stream.cuda_stream
        q = torch.zeros((2000, 3), dtype=torch.float32, device=cuda)
        ri = torch.empty((2000, 4), dtype=torch.int32, device=cuda)
        rd = torch.empty((2000, 4), dtype=torch.float32, device=cuda)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=stream):
            st = L.knn_query(t._h, q.data_ptr(), 2000, float("inf"), ri.data_ptr(), rd.data_ptr(), 4,
                             ctypes.byref(opts))
        assert st == _lib.KNN_ERR_INVALID_ARG
        assert b"knn_query_reserve" in L.knn_last_error()
        del graph
        qs = g.random((2000, 3), dtype=np.float32)
        t.batchKnn(knn.Query(torch.from_numpy(qs).to(cuda), 2000, float("inf"), ri, rd), 4)
        torch.cuda.synchronize()
        assert (ri.cpu().numpy() >= 0).all()
    finally:
        t.close()


def test_two_streams_on_one_handle_run_in_order(oracle, cuda):
    """Queries on one handle enqueued on two streams without any caller
    synchronisation between them: the second waits on the device for the
    first (they share the handle's scratch), so both answers are exact."""
    import torch
    g = np.random.Generator(np.random.Philox(4))
    pts = g.random((50000, 3), dtype=np.float32)
    t = knn.KdTreeGpu(torch.from_numpy(pts).to(cuda))
    L = _lib.load()
    try:
        s1, s2 = torch.cuda.Stream(device=cuda), torch.cuda.Stream(device=cuda)
        calls = []
        for s, m in ((s1, 30000), (s2, 20000), (s1, 25000), (s2, 30000)):
            qs = g.random((m, 3), dtype=np.float32)
            q = torch.from_numpy(qs).to(cuda)
            ri = torch.empty((m, 8), dtype=torch.int32, device=cuda)
            rd = torch.empty((m, 8), dtype=torch.float32, device=cuda)
            calls.append((s, qs, q, ri, rd))
        torch.cuda.synchronize()  # inputs uploaded; the calls below are enqueued back to back
        for s, qs, q, ri, rd in calls:
            o = _lib.QueryOpts()
            o.stream = s.cuda_stream
            _lib.check(L.knn_query(t._h, q.data_ptr(), len(qs), float("inf"), ri.data_ptr(), rd.data_ptr(), 8,
                                   ctypes.byref(o)))
        torch.cuda.synchronize()
        for i, (_, qs, _, ri, rd) in enumerate(calls):
            _same(_rows(ri, rd), oracle.bf_knn(pts, qs, 8), f"stream call {i}")
    finally:
        t.close()
