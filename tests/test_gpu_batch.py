"""GPU tests of the batched path (knn_batch_*): many uneven datasets built
and queried by two captured HIP graphs, the queries in one persistent launch.

Every dataset's result must equal the oracle's brute force bit for bit
(array-mode slot order), on every graph replay, with eager GPU work and
host synchronisations between the replays."""
import numpy as np
import pytest

import cudaexactknn_amd as knn
from cudaexactknn_amd import datasets
from tests.test_gpu_parity import dev, same

pytestmark = pytest.mark.gpu


def _make(cuda, specs, seed):
    """specs: (kind, n, dims, k, max_d2, depth_skips); queries = max(1, n // 4)
    from the same distribution (uniform in the unit cube, a few duplicates)."""
    import torch
    g = np.random.Generator(np.random.Philox(seed))
    items, host = [], []
    for kind, n, d, k, md, skips in specs:
        pts = g.random((n, d), dtype=np.float32)
        nq = max(1, n // 4)
        qs = g.random((nq, d), dtype=np.float32)
        qs[: min(8, nq)] = pts[: min(8, nq)]  # exact hits: distance 0 ties
        if n > 64:
            pts[-4:] = pts[:4]  # duplicated points: index tie-break
        p, q = dev(pts, cuda), dev(qs, cuda)
        ri = torch.full((nq, k), -7, dtype=torch.int32, device=cuda)
        rd = torch.full((nq, k), -7.0, dtype=torch.float32, device=cuda)
        items.append(knn.BatchItem(kind, p, q, k, ri, rd, md, 0, skips))
        host.append((pts, qs, k, md))
    return items, host


def _check_all(oracle, items, host, what):
    import torch
    torch.cuda.synchronize()
    for j, (it, (pts, qs, k, md)) in enumerate(zip(items, host)):
        want = oracle.bf_knn(pts, qs, k, md)
        got = (it.r_indexes.cpu().numpy(), it.r_distances.cpu().numpy())
        same(got, want, f"{what}: dataset {j} (kind {it.kind}, n {len(pts)}, d {pts.shape[1]}, k {k})")


SPECS_MIXED = [
    (knn.KIND_KD, 1000, 3, 4, np.inf, 0),
    (knn.KIND_KD, 33333, 3, 16, np.inf, 0),
    (knn.KIND_KD, 5000, 3, 32, 0.01, 0),
    (knn.KIND_KD, 2, 3, 8, np.inf, 0),
    (knn.KIND_KD, 7001, 3, 8, np.inf, 2),
    (knn.KIND_BALL, 3000, 3, 16, np.inf, 1),
    (knn.KIND_BRUTE, 2000, 16, 8, np.inf, 0),
    (knn.KIND_BRUTE, 40000, 16, 16, np.inf, 0),
    (knn.KIND_BRUTE, 777, 5, 32, np.inf, 0),
    (knn.KIND_BRUTE, 3, 16, 4, np.inf, 0),
    (knn.KIND_BRUTE, 9000, 16, 4, 0.5, 0),
]


def test_batch_mixed_replays(oracle, cuda):
    import torch
    items, host = _make(cuda, SPECS_MIXED, 41)
    print("batch: mixed kd/ball/bf16 datasets", [(s[0], s[1], s[2], s[3]) for s in SPECS_MIXED], flush=True)
    b = knn.Batch(items)
    st = b.stats()
    assert st["work_items"] > 0 and st["workgroups"] > 0
    b.run()
    _check_all(oracle, items, host, "first replay")
    # Replays: results cleared in between, eager GPU work and host syncs between them.
    for rep in range(3):
        for it in items:
            it.r_indexes.fill_(-9)
            it.r_distances.fill_(-9.0)
        x = torch.randn(512, 512, device=cuda)
        (x @ x).sum().item()
        b.run(build=(rep != 1), query=True)
        _check_all(oracle, items, host, f"replay {rep}")
    st = b.stats()
    print("batch stats:", st)
    assert st["query_ms"] > 0
    b.close()


def test_batch_separate_launch_classes(oracle, cuda):
    """Tree items with bf16 items of a wider D16 class (separate launches) and
    bf16-only D16 classes in one batch."""
    specs = [
        (knn.KIND_KD, 4096, 2, 16, np.inf, 0),
        (knn.KIND_KD, 20000, 2, 8, np.inf, 1),
        (knn.KIND_BRUTE, 3000, 40, 8, np.inf, 0),
        (knn.KIND_BRUTE, 1500, 100, 32, np.inf, 0),
        (knn.KIND_BRUTE, 2500, 20, 16, np.inf, 0),
    ]
    items, host = _make(cuda, specs, 42)
    b = knn.Batch(items, build_streams=2)
    b.run()
    _check_all(oracle, items, host, "separate launches")
    b.close()


def test_batch_matches_single_structures(oracle, cuda):
    """A slice of the uneven 256-dataset workload (seed 4), smallest sets,
    against the single-dataset path (kd tree / bf16 brute force)."""
    import torch
    sets = datasets.uneven_batch(24, 1024, 60000, seed=4)
    items = []
    for s in sets:
        p, q = dev(s["points"], cuda), dev(s["queries"], cuda)
        nq, k = len(s["queries"]), s["k"]
        kind = knn.KIND_KD if s["dims"] == 3 else knn.KIND_BRUTE
        items.append(knn.BatchItem(kind, p, q, k, torch.empty((nq, k), dtype=torch.int32, device=cuda),
                                   torch.empty((nq, k), dtype=torch.float32, device=cuda)))
    b = knn.Batch(items)
    b.run()
    torch.cuda.synchronize()
    for j, (s, it) in enumerate(zip(sets, items)):
        nq, k = len(s["queries"]), s["k"]
        if it.kind == knn.KIND_KD:
            t = knn.KdTreeGpu(it.points, depth_skips=0)
            form = knn.FORM_PLAIN
        else:
            t = knn.BruteForceGpu(it.points)
            form = knn.FORM_BF16
        ri = torch.empty((nq, k), dtype=torch.int32, device=cuda)
        rd = torch.empty((nq, k), dtype=torch.float32, device=cuda)
        t.batchKnn(knn.Query(it.queries, nq, float("inf"), ri, rd), k, form=form)
        torch.cuda.synchronize()
        same((it.r_indexes.cpu().numpy(), it.r_distances.cpu().numpy()), (ri.cpu().numpy(), rd.cpu().numpy()),
             f"dataset {j} vs single structure")
        if j % 6 == 0:
            want = oracle.bf_knn(s["points"], s["queries"], k)
            same((ri.cpu().numpy(), rd.cpu().numpy()), want, f"dataset {j} vs oracle")
        t.close()
    b.close()


def test_batch_rejects_bad_items(cuda):
    import torch
    p = torch.rand((100, 3), device=cuda)
    q = torch.rand((10, 3), device=cuda)
    ri = torch.empty((10, 64), dtype=torch.int32, device=cuda)
    rd = torch.empty((10, 64), dtype=torch.float32, device=cuda)
    with pytest.raises(knn.KnnError):
        knn.Batch([knn.BatchItem(knn.KIND_KD, p, q, 64, ri, rd)])  # k above 32
    p5 = torch.rand((100, 5), device=cuda)
    q5 = torch.rand((10, 5), device=cuda)
    with pytest.raises(knn.KnnError):
        knn.Batch([knn.BatchItem(knn.KIND_KD, p5, q5, 4, ri, rd)])  # tree dims 5
    with pytest.raises(knn.KnnError):
        knn.Batch([knn.BatchItem(knn.KIND_KD, p, q, 4, ri, rd),
                   knn.BatchItem(knn.KIND_KD, torch.rand((100, 2), device=cuda), torch.rand((10, 2), device=cuda),
                                 4, ri, rd)])  # mixed tree dims
