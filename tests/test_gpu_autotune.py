"""GPU tests of the autotuner: every candidate option set returns the same
result (the tuner raises otherwise), the winner is persisted per
(kind, D, K, size bucket) and reused, and tuned options stay exact."""
import numpy as np
import pytest

import cudaexactknn_amd as knn
from cudaexactknn_amd import autotune
from tests.test_gpu_parity import dev, run, same

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,d,k", [("kd", 3, 8), ("ball", 3, 16), ("brute", 16, 16)])
def test_tuner_candidates_agree_and_persist(oracle, cuda, tmp_path, kind, d, k):
    g = np.random.Generator(np.random.Philox(77 + d + k))
    pts = g.random((30000, d), dtype=np.float32)
    qs = g.random((3000, d), dtype=np.float32)
    qs[:5] = pts[:5]
    p = dev(pts, cuda)
    if kind == "kd":
        s = knn.KdTreeGpu(p, depth_skips=0)
    elif kind == "ball":
        s = knn.BallTreeGpu(p, depth_skips=1)
    else:
        s = knn.BruteForceGpu(p)
    path = str(tmp_path / "tuned.json")
    print(f"autotune {kind} D={d} K={k}", flush=True)
    r = autotune.tune_and_store(s, dev(qs, cuda), k, path=path, repeats=1)
    assert len(r["results"]) >= 8
    table = autotune.load_table(path)
    key = autotune.table_key(kind, d, k, 30000)
    assert key in table and table[key]["best"] == r["best"]
    opts = autotune.tuned_options(s, k, path=path)
    got = run(s, qs, k, cuda, **opts)
    want = oracle.bf_knn(pts, qs, k)
    if opts.get("USE_HEAP_TRACKING"):
        got = (got[0][:, ::-1].copy(), got[1][:, ::-1].copy())
    same(got, want, f"tuned {kind} options {opts}")
    s.close()


def test_tune_batch_options_agree(oracle, cuda):
    import torch
    g = np.random.Generator(np.random.Philox(5))
    items, host = [], []
    for kind, n, d, k in [(knn.KIND_KD, 20000, 3, 8), (knn.KIND_BRUTE, 8000, 16, 16), (knn.KIND_KD, 3000, 3, 4)]:
        pts = g.random((n, d), dtype=np.float32)
        qs = g.random((n // 4, d), dtype=np.float32)
        items.append(knn.BatchItem(kind, dev(pts, cuda), dev(qs, cuda), k,
                                   torch.empty((n // 4, k), dtype=torch.int32, device=cuda),
                                   torch.empty((n // 4, k), dtype=torch.float32, device=cuda)))
        host.append((pts, qs, k))
    r = autotune.tune_batch(items, cands=[dict(blocks_per_cu=0, tile=0), dict(blocks_per_cu=1, tile=1),
                                          dict(blocks_per_cu=2, tile=4, leaf_span=0xFFFFFFFF)])
    assert len(r["results"]) == 3
    torch.cuda.synchronize()
    for it, (pts, qs, k) in zip(items, host):
        same((it.r_indexes.cpu().numpy(), it.r_distances.cpu().numpy()), oracle.bf_knn(pts, qs, k), "batch")
