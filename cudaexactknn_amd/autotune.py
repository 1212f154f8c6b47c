"""Autotuner for batchKnn launch options.

For one structure and a representative query set it times every candidate
option set (BLOCK_SIZE, K-buffer kind = USE_HEAP_TRACKING x
USE_TRACKING_SHMEM, USE_STACK_SHMEM, traversal shape, leaf span, bf16 tile
staging, form) with the library's hipEvent timing, checks that every
candidate returns the same result (heap-mode rows compared in reverse, as
they are emitted descending), and keeps the fastest.  Winners are stored in a
JSON table keyed by (kind, D, K, size bucket = ceil(log2 n)) so later calls
look them up instead of re-tuning.

Options never change results: the tuner raises if two candidates disagree.
"""
from __future__ import annotations

import itertools
import json
import math
import os
from typing import Any, Iterable

from . import FORM_BF16, FORM_PLAIN, BallTreeGpu, BruteForceGpu, KdTreeGpu, Query, KnnError
from . import _lib

DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_options.json")


def _kind_name(struct) -> str:
    if isinstance(struct, KdTreeGpu):
        return "kd"
    if isinstance(struct, BallTreeGpu):
        return "ball"
    if isinstance(struct, BruteForceGpu):
        return "brute"
    raise TypeError(f"not a knn structure: {type(struct)!r}")


def table_key(kind: str, dims: int, k: int, n: int) -> str:
    return f"{kind}:D{dims}:K{k}:N2^{max(0, math.ceil(math.log2(max(1, n))))}"


def candidates(kind: str, dims: int, k: int) -> list[dict]:
    """Candidate option sets for a structure kind (all are exact)."""
    out = []
    if kind == "brute":
        for bs, heap, lds in itertools.product((64, 128, 256), (False, True), (False, True)):
            out.append(dict(BLOCK_SIZE=bs, USE_HEAP_TRACKING=heap, USE_TRACKING_SHMEM=lds, form=FORM_PLAIN))
        if k <= 32 and dims <= 256:
            for heap, tile in itertools.product((False, True), (0, 1, 2, 4, 8, 16, 32)):
                out.append(dict(BLOCK_SIZE=256, USE_HEAP_TRACKING=heap, USE_TRACKING_SHMEM=False, form=FORM_BF16,
                                tile=tile))
        return out
    for bs, heap, span in itertools.product((64, 128, 256), (False, True), (0, -1, 1)):
        out.append(dict(BLOCK_SIZE=bs, USE_HEAP_TRACKING=heap, USE_TRACKING_SHMEM=False, USE_STACK_SHMEM=False,
                        per_lane=False, leaf_span_delta=span))
    for bs, heap, lds, stack in itertools.product((64, 128, 256), (False, True), (False, True), (False, True)):
        out.append(dict(BLOCK_SIZE=bs, USE_HEAP_TRACKING=heap, USE_TRACKING_SHMEM=lds, USE_STACK_SHMEM=stack,
                        per_lane=True, leaf_span_delta=0))
    return out


def _default_span(kind: str, info: dict, k: int, per_lane: bool) -> int:
    """The library's default leaf span (knn_capi.hip default_span)."""
    target = k // 2 if (kind == "ball" and info["dims"] >= 8 and k >= 2) else k
    mc = max(1, info["max_children"])
    s = 0 if target <= mc else math.ceil(math.log2((target + mc - 1) // mc))
    return min(s, info["node_depth"])


def _resolve(kind: str, struct, k: int, cand: dict) -> dict:
    """Candidate -> batchKnn keyword arguments (leaf span made absolute)."""
    opts = {key: v for key, v in cand.items() if key != "leaf_span_delta"}
    if kind != "brute":
        base = _default_span(kind, struct.info(), k, cand.get("per_lane", False))
        span = base + cand.get("leaf_span_delta", 0)
        span = max(0, min(span, struct.info()["node_depth"]))
        opts["leaf_span"] = 0xFFFFFFFF if span == 0 else span
    return opts


def tune(struct, queries: Any, k: int, max_distance2: float = float("inf"),
         cands: Iterable[dict] | None = None, repeats: int = 3) -> dict:
    """Time every candidate on `queries` (a device tensor [q][D]) and return
    {"best": options, "ms": best time, "results": [(options, ms), ...]}.
    Raises KnnError if two candidates return different results."""
    import torch

    kind = _kind_name(struct)
    dims = int(queries.shape[1])
    nq = int(queries.shape[0])
    cands = list(cands) if cands is not None else candidates(kind, dims, k)
    dev = queries.device
    ri = torch.empty((nq, k), dtype=torch.int32, device=dev)
    rd = torch.empty((nq, k), dtype=torch.float32, device=dev)
    q = Query(queries, nq, max_distance2, ri, rd)
    ref = None
    results = []
    for cand in cands:
        opts = _resolve(kind, struct, k, cand)
        try:
            struct.batchKnn(q, k, **opts)
        except KnnError as e:
            if e.status in (_lib.KNN_ERR_UNSUPPORTED, _lib.KNN_ERR_INVALID_ARG):
                continue  # e.g. an LDS request above the workgroup limit
            raise
        torch.cuda.synchronize()
        got_i, got_d = ri.clone(), rd.clone()
        if opts.get("USE_HEAP_TRACKING"):
            got_i, got_d = got_i.flip(1), got_d.flip(1)
        if ref is None:
            ref = (got_i, got_d)
        elif not (torch.equal(ref[0], got_i) and torch.equal(ref[1].view(torch.int32), got_d.view(torch.int32))):
            raise KnnError(_lib.KNN_ERR_INTERNAL, f"option set {cand} changed the result")
        times = []
        for _ in range(repeats):
            struct.batchKnn(q, k, **opts)
            times.append(struct.last_query_stats()["query_ms"])
        results.append((cand, min(times)))
    if not results:
        raise KnnError(_lib.KNN_ERR_UNSUPPORTED, "no candidate option set ran")
    best = min(results, key=lambda r: r[1])
    return {"best": best[0], "ms": best[1], "results": results}


def load_table(path: str = DEFAULT_TABLE) -> dict:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def save_table(table: dict, path: str = DEFAULT_TABLE) -> None:
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tuned_options(struct, k: int, table: dict | None = None, path: str = DEFAULT_TABLE) -> dict:
    """batchKnn keyword arguments for `struct` from the table ({} = library
    defaults when the bucket was never tuned)."""
    kind = _kind_name(struct)
    table = load_table(path) if table is None else table
    entry = table.get(table_key(kind, struct.dims, k, struct.size))
    return _resolve(kind, struct, k, entry["best"]) if entry else {}


def tune_and_store(struct, queries: Any, k: int, max_distance2: float = float("inf"),
                   path: str = DEFAULT_TABLE, **kw) -> dict:
    """Tune and record the winner under (kind, D, K, size bucket)."""
    r = tune(struct, queries, k, max_distance2, **kw)
    table = load_table(path)
    table[table_key(_kind_name(struct), struct.dims, k, struct.size)] = {
        "best": r["best"], "ms": r["ms"], "queries": int(queries.shape[0]), "candidates": len(r["results"])}
    save_table(table, path)
    return r


def tune_batch(items: list, cands: Iterable[dict] | None = None, repeats: int = 2) -> dict:
    """Time a Batch under each batch-level option set (persistent workgroups
    per CU, bf16 tile staging, leaf span) and return the fastest; results must
    agree across all of them."""
    import torch

    from . import Batch

    if cands is None:
        cands = [dict(blocks_per_cu=b, tile=t) for b in (0, 2, 4) for t in (0, 2, 4)]
    ref, results = None, []
    for cand in cands:
        b = Batch(items, **cand)
        b.run()
        torch.cuda.synchronize()
        out = [(it.r_indexes.clone(), it.r_distances.clone()) for it in items]
        if ref is None:
            ref = out
        else:
            for (a_i, a_d), (b_i, b_d) in zip(ref, out):
                if not (torch.equal(a_i, b_i) and torch.equal(a_d.view(torch.int32), b_d.view(torch.int32))):
                    raise KnnError(_lib.KNN_ERR_INTERNAL, f"batch option set {cand} changed the result")
        times = []
        for _ in range(repeats):
            b.run(build=False)
            times.append(b.stats()["query_ms"])
        results.append((cand, min(times)))
        b.close()
    best = min(results, key=lambda r: r[1])
    return {"best": best[0], "ms": best[1], "results": results}
