"""GPU parity on the measured workloads' own data and sizes.

- tree sizing of the library against the hand-written table of
  tests/pins.py (src/KdTree.cuh:56-62, src/BallTree.cuh:123-128);
- the 32-Gaussian 3-D mixture (KdTreeGpu<3,3,256> depthSkips 0 and
  BallTreeGpu<3,256> depthSkips 1, K = 16, seed 1) at 2^20 points against
  the oracle's brute force and tree exports, and at its full size of 2^24
  points / 2^22 queries on a sampled query subset plus size-independent
  checks (kd == ball on every query, sorted rows, full kd structure);
- the bf16-filtered brute force at sizes where its seeding pass runs
  (n >= 32,768 points, D > 16), against the oracle's brute force.
"""
import numpy as np
import pytest

import cudaexactknn_amd as knn
from cudaexactknn_amd import datasets
from tests import pins
from tests.test_gpu_parity import dev, run, same

pytestmark = pytest.mark.gpu


def test_library_tree_sizing_matches_table(cuda):
    g = np.random.Generator(np.random.Philox(12))
    for (n, skips) in pins.SIZING_TABLE:
        if n > 65537:
            continue
        p = dev(g.random((n, 3), dtype=np.float32), cuda)
        for ball in (False, True):
            t = knn.BallTreeGpu(p, depth_skips=skips) if ball else knn.KdTreeGpu(p, depth_skips=skips)
            ok, msg = pins.check_sizing(t.info(), n, skips, ball)
            assert ok, msg
            t.close()


def _rows_sorted(ri, rd):
    """Array-mode rows are ascending by (distance bits, index)."""
    keys = (rd.view(np.uint32).astype(np.uint64) << np.uint64(32)) | ri.view(np.uint32).astype(np.uint64)
    return bool(np.all(keys[:, 1:] > keys[:, :-1]))


def test_gaussian_mixture_2p20_against_oracle(oracle, cuda):
    """2^20 points and 2^14 queries of the 32-Gaussian mixture recipe:
    every row of both trees, per-lane and wave-cooperative, against the
    oracle's brute force; the kd export bit for bit against the oracle's kd
    build; the ball export's dims and leaves exactly and centroids within
    2 ulp against the oracle's ball build."""
    pts, qs = datasets.gaussian_mixture_3d(1 << 20, 1 << 14, seed=1)
    want = oracle.bf_knn(pts, qs, 16)
    want_h = (want[0][:, ::-1].copy(), want[1][:, ::-1].copy())
    p = dev(pts, cuda)
    kd = knn.KdTreeGpu(p, depth_skips=0, split_dimensions=3, block_size=256)
    ball = knn.BallTreeGpu(p, depth_skips=1, block_size=256)
    for name, t in (("kd", kd), ("ball", ball)):
        for per_lane in (False, True):
            same(run(t, qs, 16, cuda, BLOCK_SIZE=256 if per_lane else 0, per_lane=per_lane), want,
                 f"{name} lane={per_lane}")
        same(run(t, qs, 16, cuda, USE_HEAP_TRACKING=True), want_h, f"{name} heap")
    okd = oracle.KdTree(pts, 0, 3)
    nodes, leaf_idx, leaf_pts, _ = kd.export()
    onodes, oleaf = okd.export()
    assert np.array_equal(nodes.view(np.uint32), onodes.view(np.uint32))
    assert np.array_equal(leaf_idx, oleaf)
    m = leaf_idx >= 0
    assert np.array_equal(leaf_pts[m], pts[leaf_idx[m]])
    obl = oracle.BallTree(pts, 1)
    bnodes, bleaf, _, bdims = ball.export()
    onodes, odims, oleaf = obl.export()
    assert np.array_equal(bdims, odims)
    assert np.array_equal(bleaf, oleaf)
    a, b = bnodes.reshape(onodes.shape).astype(np.float64), onodes.astype(np.float64)
    assert np.all(np.abs(a - b) <= 2 * np.spacing(np.abs(onodes)).astype(np.float64))


def test_gaussian_mixture_full_size(oracle, cuda):
    """The real-size configuration: 16,777,216 points and 4,194,304 queries.
    Sizing from the hand-written table; the whole kd export against the
    oracle's kd build; every query row of the kd tree equal to the ball
    tree's and sorted; 4,096 sampled queries against the oracle's kd tree
    (itself pinned to brute force)."""
    import torch
    pts, qs = datasets.gaussian_mixture_3d(1 << 24, 1 << 22, seed=1)
    p = dev(pts, cuda)
    kd = knn.KdTreeGpu(p, depth_skips=0, split_dimensions=3, block_size=256)
    ok, msg = pins.check_sizing(kd.info(), 1 << 24, 0, False)
    assert ok, msg
    okd = oracle.KdTree(pts, 0, 3)
    nodes, leaf_idx, _, _ = kd.export()
    onodes, oleaf = okd.export()
    assert np.array_equal(nodes.view(np.uint32), onodes.view(np.uint32))
    assert np.array_equal(leaf_idx, oleaf)
    del nodes, leaf_idx, onodes, oleaf
    q = dev(qs, cuda)
    ri = torch.empty((len(qs), 16), dtype=torch.int32, device=cuda)
    rd = torch.empty((len(qs), 16), dtype=torch.float32, device=cuda)
    kd.batchKnn(knn.Query(q, len(qs), float("inf"), ri, rd), 16)
    torch.cuda.synchronize()
    kd_i, kd_d = ri.cpu().numpy(), rd.cpu().numpy()
    kd.close()
    ball = knn.BallTreeGpu(p, depth_skips=1, block_size=256)
    ok, msg = pins.check_sizing(ball.info(), 1 << 24, 1, True)
    assert ok, msg
    ball.batchKnn(knn.Query(q, len(qs), float("inf"), ri, rd), 16)
    torch.cuda.synchronize()
    same((ri.cpu().numpy(), rd.cpu().numpy()), (kd_i, kd_d), "kd vs ball, every query")
    assert _rows_sorted(kd_i, kd_d)
    sample = np.random.Generator(np.random.Philox(101)).choice(len(qs), 4096, replace=False)
    want = okd.knn(qs[sample], 16)
    same((kd_i[sample], kd_d[sample]), want, "sampled rows vs oracle kd")


@pytest.mark.parametrize("d", [32, 64, 128])
def test_bf16_seeded_screen_against_oracle(oracle, cuda, d):
    """65,536 scaled-normal points (2,048 point tiles: the seeding pass runs
    for D > 16), 512 queries, K = 32: array and heap rows, unbounded and
    finite radius, equal to the oracle's brute force; the query statistics
    confirm the seeding pass ran."""
    pts, qs = datasets.scaled_normal(1 << 16, 512, d, seed=2)
    qs[0] = pts[17]
    pts[300] = pts[301]
    qs[1] = pts[301]
    bf = knn.BruteForceGpu(dev(pts, cuda))
    want_inf = oracle.bf_knn(pts, qs, 32)
    md = float(np.float32(np.median(want_inf[1][:, 20])))
    for radius in (np.inf, md):
        for heap in (False, True):
            want = oracle.bf_knn(pts, qs, 32, max_d2=radius, heap=heap)
            got = run(bf, qs, 32, cuda, md=radius, USE_HEAP_TRACKING=heap, form=knn.FORM_BF16)
            same(got, want, f"bf16 seeded d={d} heap={heap} md={radius}")
            st = bf.last_query_stats()
            assert st["bf16_seeded"] == 1.0, st
    for k in (1, 5, 16):
        same(run(bf, qs, k, cuda, form=knn.FORM_BF16), oracle.bf_knn(pts, qs, k), f"bf16 seeded d={d} k={k}")
