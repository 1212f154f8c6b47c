"""Host-resident queries (knn_query_host): chunked, with uploads and
downloads overlapping the search; results identical to the device path."""
import numpy as np
import pytest

import cudaexactknn_amd as knn
from tests.test_gpu_parity import dev, run, same

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["kd", "ball", "brute", "brute_bf16"])
@pytest.mark.parametrize("chunk,pinned", [(3000, True), (0, False), (777, False)])
def test_host_pipeline_matches_device(oracle, cuda, kind, chunk, pinned):
    g = np.random.Generator(np.random.Philox(123))
    d = 16 if kind.startswith("brute") else 3
    pts = g.random((40000, d), dtype=np.float32)
    qs = g.random((10001, d), dtype=np.float32)
    p = dev(pts, cuda)
    s = {"kd": lambda: knn.KdTreeGpu(p, depth_skips=0), "ball": lambda: knn.BallTreeGpu(p, depth_skips=1),
         "brute": lambda: knn.BruteForceGpu(p), "brute_bf16": lambda: knn.BruteForceGpu(p)}[kind]()
    opts = {"form": knn.FORM_BF16} if kind == "brute_bf16" else {}
    K = 8
    print(f"host pipeline {kind} chunk {chunk} pinned {pinned}", flush=True)
    if pinned:
        pq = knn.pinned_array(qs.shape, np.float32)
        pq.array[:] = qs
        pi = knn.pinned_array((len(qs), K), np.int32)
        pd = knn.pinned_array((len(qs), K), np.float32)
        got = s.batchKnnHost(pq.array, K, r_indexes=pi.array, r_distances=pd.array, chunk=chunk, **opts)
        got = (got[0].copy(), got[1].copy())
        for a in (pq, pi, pd):
            a.close()
    else:
        got = s.batchKnnHost(qs, K, chunk=chunk, **opts)
    same(got, run(s, qs, K, cuda, **opts), f"{kind} host pipeline vs device path")
    if kind == "kd":
        same(got, oracle.bf_knn(pts, qs, K), "kd host pipeline vs oracle")
    assert s.last_query_stats()["query_ms"] >= 0
    s.close()


def test_host_pipeline_empty_and_errors(cuda):
    g = np.random.Generator(np.random.Philox(5))
    p = dev(g.random((1000, 3), dtype=np.float32), cuda)
    t = knn.KdTreeGpu(p)
    i, d = t.batchKnnHost(np.zeros((0, 3), np.float32), 4)
    assert i.shape == (0, 4)
    with pytest.raises(knn.KnnError):
        t.batchKnnHost(np.zeros((5, 2), np.float32), 4)
