"""GPU tests of the autotuner and of its table inside the library.

- every candidate option set returns the same rows (the tuner raises
  otherwise), the winner is stored per (class, D, K, size bucket) and tuned
  options stay exact against the oracle;
- the table compiled into the library is what knn_query runs with when the
  caller leaves BLOCK_SIZE / leaf span / tile at 0 (knn_query_resolved_opts),
  and those defaults return the oracle's rows;
- batch-level option sets agree.
"""
import ctypes

import numpy as np
import pytest

import cudaexactknn_amd as knn
from cudaexactknn_amd import _lib, autotune
from tests.test_gpu_parity import dev, run, same

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cls,d,k,heap", [("kd", 3, 8, False), ("ball", 3, 16, True), ("brute", 16, 16, False),
                                          ("bf16", 32, 16, False)])
def test_tuner_candidates_agree_and_persist(oracle, cuda, tmp_path, cls, d, k, heap):
    g = np.random.Generator(np.random.Philox(77 + d + k))
    pts = g.random((30000, d), dtype=np.float32)
    qs = g.random((3000, d), dtype=np.float32)
    qs[:5] = pts[:5]
    p = dev(pts, cuda)
    if cls == "kd":
        s = knn.KdTreeGpu(p, depth_skips=0)
    elif cls == "ball":
        s = knn.BallTreeGpu(p, depth_skips=1)
    else:
        s = knn.BruteForceGpu(p)
    form = knn.FORM_BF16 if cls == "bf16" else knn.FORM_PLAIN
    path = str(tmp_path / "tuned.json")
    print(f"autotune {cls} D={d} K={k}", flush=True)
    base = {"USE_HEAP_TRACKING": True} if heap else {}
    r = autotune.tune_and_store(s, dev(qs, cuda), k, form=form, base=base, path=path, repeats=1)
    assert len(r["results"]) >= 4
    table = autotune.load_table(path)
    key = autotune.table_key(cls, d, k, 30000)
    assert key in table and table[key]["best"] == r["best"]
    opts = autotune.tuned_options(s, k, form=form, path=path)
    got = run(s, qs, k, cuda, USE_HEAP_TRACKING=heap, **opts)
    want = oracle.bf_knn(pts, qs, k, heap=heap)
    same(got, want, f"tuned {cls} options {opts}")
    s.close()


def _committed_entries(cls, dims):
    table = autotune.load_table()
    out = []
    for key, e in table.items():
        c, d, k, nb = key.split(":")
        if c == cls and int(d[1:]) == dims:
            out.append((int(k[1:]), int(nb[3:]), e["best"]))
    return sorted(out)


@pytest.mark.parametrize("cls,dims", [("kd", 3), ("bf16", 16)])
def test_library_defaults_follow_the_compiled_table(oracle, cuda, cls, dims):
    """For committed entries of the batched workload's grid: the options a
    default knn_query resolves to are the table's, and its rows equal the
    oracle's and those of the explicit built-in defaults."""
    entries = [e for e in _committed_entries(cls, dims) if 12 <= e[1] <= 15]
    if not entries:
        pytest.skip(f"no committed {cls} D={dims} entries between 2^12 and 2^15 points")
    L = _lib.load()
    for k, nb, best in entries[:4]:
        n = 1 << nb
        g = np.random.Generator(np.random.Philox(n + k))
        pts = g.random((n, dims), dtype=np.float32)
        qs = g.random((n // 4, dims), dtype=np.float32)
        p = dev(pts, cuda)
        s = knn.KdTreeGpu(p, depth_skips=0) if cls == "kd" else knn.BruteForceGpu(p)
        form = knn.FORM_BF16 if cls == "bf16" else knn.FORM_PLAIN
        lib_entry = autotune.library_lookup(cls, dims, k, n)
        assert lib_entry is not None, f"{cls} D={dims} K={k} n=2^{nb}: entry missing from the library"
        assert lib_entry["BLOCK_SIZE"] == int(best.get("BLOCK_SIZE", 0) or 0)
        assert lib_entry["leaf_span_delta"] == int(best.get("leaf_span_delta", 0) or 0)
        assert lib_entry["tile"] == int(best.get("tile", 0) or 0)
        o = _lib.QueryOpts()
        o.form = form
        out = _lib.QueryOpts()
        _lib.check(L.knn_query_resolved_opts(s.handle, len(qs), k, ctypes.byref(o), ctypes.byref(out)))
        want_opts = autotune._resolve(cls, s, k, best)
        assert out.block == want_opts["BLOCK_SIZE"], (out.block, want_opts)
        if cls == "kd":
            assert out.leaf_span == want_opts["leaf_span"], (out.leaf_span, want_opts)
        else:
            assert out.tile == int(best.get("tile", 0) or 0), (out.tile, best)
        want = oracle.bf_knn(pts, qs, k)
        same(run(s, qs, k, cuda, form=form), want, f"{cls} K={k} n=2^{nb} tuned defaults")
        builtin = autotune._resolve(cls, s, k, autotune.candidates(cls, dims, k)[0])
        same(run(s, qs, k, cuda, **builtin), want, f"{cls} K={k} n=2^{nb} built-in defaults")
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
