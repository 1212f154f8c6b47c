"""The reference's speed-test sweeps on seeded inputs, for continuity with
its workload shapes (GPU structures only; one JSON line per measurement).

Shapes and template arguments follow the reference's drivers:
  speedup       tests/SpeedupSpeedTests.cu: D=5, K=10, maxDistance2 5.9;
                query counts 1k..100k on 100k points, point counts 10k..1M
                with 10k queries.  KdTreeGpu<D,D,32>(depthSkips 3)
                batchKnn<K,32,false,false,true>, BallTreeGpu<D,32>(1)
                batchKnn<K,16,true,true,true>, BruteForceGpu<D>
                batchKnn<K,32,false,false>.
  construction  tests/ConstructionSpeedTests.cu: build times, D=3 with
                100k..2.5M points, then D in {3,5,7,10,15,20} at 500k.
  dim3          tests/Dim3SpeedTests.cu: D=3, K=10, maxDistance2 5.9;
                250k..10M queries on 1M points, 250k..10M points with 1M
                queries.
  dimx          tests/DimXSpeedTests.cu: D=6..17, 500k points, 10k queries,
                K=5, maxDistance2 0.01 * D (DimXSpeedTests.cu:19).

Inputs are uniform in [0, 1)^D from a seeded Philox stream (the reference
drew them from std::random_device).  Timings are the library's hipEvent
figures: query_ms per batchKnn (mean of 5 after one warm-up, as the
reference averaged 5 of 6) and build_ms per construction.  The reference
also timed its CPU classes in the speedup sweep; their CPU counterparts are
the header-only host classes of include/KnnTrees.hpp, outside this
Python-driven GPU sweep.

    python -m cudaexactknn_amd.speed_suite speedup [--quick]
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

INF = float("inf")


def _points(n: int, d: int, seed: int) -> np.ndarray:
    g = np.random.Generator(np.random.Philox(seed))
    return g.random((n, d), dtype=np.float32)


def _time_query(struct, q, K, max_d2, reps, **opts):
    """Mean query ms over `reps` calls after one warm-up, and the result rows."""
    import torch

    import cudaexactknn_amd as knn
    nq = q.shape[0]
    ri = torch.empty((nq, K), dtype=torch.int32, device=q.device)
    rd = torch.empty((nq, K), dtype=torch.float32, device=q.device)
    Q = knn.Query(q, nq, max_d2, ri, rd)
    struct.batchKnn(Q, K, **opts)
    ms = []
    for _ in range(reps):
        struct.batchKnn(Q, K, **opts)
        ms.append(struct.last_query_stats()["query_ms"])
    return float(np.mean(ms)), (ri, rd)


def _structures(p, d):
    import cudaexactknn_amd as knn
    return [
        ("KdTreeGpu", lambda: knn.KdTreeGpu(p, depth_skips=3, split_dimensions=d, block_size=32),
         dict(BLOCK_SIZE=32, USE_HEAP_TRACKING=False, USE_TRACKING_SHMEM=False, USE_STACK_SHMEM=True)),
        ("BallTreeGpu", lambda: knn.BallTreeGpu(p, depth_skips=1, block_size=32),
         dict(BLOCK_SIZE=16, USE_HEAP_TRACKING=True, USE_TRACKING_SHMEM=True, USE_STACK_SHMEM=True)),
        ("BruteForceGpu", lambda: knn.BruteForceGpu(p),
         dict(BLOCK_SIZE=32, USE_HEAP_TRACKING=False, USE_TRACKING_SHMEM=False)),
    ]


def _emit(rec: dict) -> None:
    print(json.dumps(rec), flush=True)


def _query_sweep(suite, d, K, max_d2, n_pts_list, n_qry_list, vary, reps, seed, kinds=None, on_result=None):
    """Times each structure at each (points, queries) size.  `on_result`, if
    given, receives (record, points, queries, (indices, distances)) after
    each measurement, for result checks outside the timed calls."""
    import torch
    dev = torch.device("cuda:0")
    for n, nq in zip(n_pts_list, n_qry_list):
        pn, qn = _points(n, d, seed), _points(nq, d, seed + 1)
        p = torch.from_numpy(pn).to(dev)
        q = torch.from_numpy(qn).to(dev)
        for name, make, opts in _structures(p, d):
            if kinds and name not in kinds:
                continue
            s = make()
            ms, rows = _time_query(s, q, K, max_d2, reps, **opts)
            rec = {"suite": suite, "structure": name, "dims": d, "points": n, "queries": nq, "k": K,
                   "max_distance2": max_d2, "vary": vary, "query_ms": round(ms, 4),
                   "queries_per_s": round(nq / ms * 1e3, 1)}
            _emit(rec)
            if on_result:
                on_result(rec, pn, qn, rows)
            s.close()


def run_speedup(quick: bool, reps: int, on_result=None) -> None:
    q_counts = [1000, 5000, 10000, 25000, 50000, 100000]
    p_counts = [10000, 50000, 100000, 250000, 500000, 1000000]
    if quick:
        q_counts, p_counts = q_counts[:2], p_counts[:2]
    _query_sweep("speedup", 5, 10, 5.9, [100000] * len(q_counts), q_counts, "queries", reps, 10, None, on_result)
    _query_sweep("speedup", 5, 10, 5.9, p_counts, [10000] * len(p_counts), "points", reps, 20, None, on_result)


def run_construction(quick: bool, reps: int, on_result=None) -> None:
    import torch

    import cudaexactknn_amd as knn
    dev = torch.device("cuda:0")
    cases = [(n, 3) for n in (100000, 250000, 500000, 1000000, 2500000)]
    cases += [(500000, d) for d in (3, 5, 7, 10, 15, 20)]
    if quick:
        cases = cases[:2] + cases[5:7]
    for n, d in cases:
        p = torch.from_numpy(_points(n, d, 30 + d)).to(dev)
        for name, make in (("KdTreeGpu", lambda: knn.KdTreeGpu(p, depth_skips=3, split_dimensions=d, block_size=32)),
                           ("BallTreeGpu", lambda: knn.BallTreeGpu(p, depth_skips=1, block_size=32))):
            make().close()
            ms = []
            for _ in range(reps):
                s = make()
                ms.append(s.build_ms())
                s.close()
            _emit({"suite": "construction", "structure": name, "dims": d, "points": n,
                   "build_ms": round(float(np.mean(ms)), 4)})


def run_dim3(quick: bool, reps: int, on_result=None) -> None:
    counts = [250000, 500000, 1000000, 2500000, 5000000, 10000000]
    if quick:
        counts = counts[:2]
    kinds = ("KdTreeGpu", "BallTreeGpu")
    _query_sweep("dim3", 3, 10, 5.9, [1000000] * len(counts), counts, "queries", reps, 40, kinds, on_result)
    _query_sweep("dim3", 3, 10, 5.9, counts, [1000000] * len(counts), "points", reps, 50, kinds, on_result)


def run_dimx(quick: bool, reps: int, on_result=None) -> None:
    dims = list(range(6, 18))
    if quick:
        dims = dims[:2]
    for d in dims:
        max_d2 = float(np.float32(0.01 * d))
        _query_sweep("dimx", d, 5, max_d2, [500000], [10000], "dims", reps, 60 + d, None, on_result)


SUITES = {"speedup": run_speedup, "construction": run_construction, "dim3": run_dim3, "dimx": run_dimx}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("suite", choices=sorted(SUITES) + ["all"])
    ap.add_argument("--quick", action="store_true", help="first sizes of each sweep only")
    ap.add_argument("--reps", type=int, default=5)
    a = ap.parse_args(argv)
    for name in (sorted(SUITES) if a.suite == "all" else [a.suite]):
        SUITES[name](a.quick, a.reps)
    return 0


if __name__ == "__main__":
    sys.exit(main())
