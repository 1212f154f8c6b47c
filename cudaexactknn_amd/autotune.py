"""Autotuner for batchKnn launch options.

For one structure and a representative query set it times every candidate
set of result-neutral launch options -- BLOCK_SIZE, the leaf span of the
wave-cooperative tree traversal (relative to the library's default), the
bf16 form's tile staging -- with the library's hipEvent timing, checks that
every candidate returns the same rows, and keeps the fastest.  The caller's
semantic options (USE_HEAP_TRACKING, USE_TRACKING_SHMEM, USE_STACK_SHMEM, the
form) are held fixed.  Winners are stored in a JSON table keyed by
(class, D, K, size bucket = ceil(log2 n)), class one of "brute" (plain form),
"bf16" (bf16 form), "kd", "ball".

The library compiles the committed table (tuned_options.json next to this
file) into knn_query's and knn_batch_create's defaults: a zero BLOCK_SIZE,
leaf span or tile takes the table's choice for the nearest tuned size within
one bucket.  `python -m cudaexactknn_amd.autotune --sweep` re-measures the
table on the workload shapes the library is benchmarked on.
"""
from __future__ import annotations

import argparse
import ctypes
import itertools
import json
import math
import os
import sys
import time
from typing import Any, Iterable

from . import FORM_BF16, FORM_PLAIN, BallTreeGpu, BruteForceGpu, KdTreeGpu, Query, KnnError
from . import _lib

DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_options.json")
CLASSES = {"brute": 0, "kd": 1, "ball": 2, "bf16": 3}
# Built-in defaults of the library (knn_capi.hip default_block): what a query
# runs with when the table has no entry.
BUILTIN_BLOCK = {"brute": 256, "bf16": 256, "kd": 64, "ball": 64}


def class_of(struct, form: int = FORM_PLAIN) -> str:
    if isinstance(struct, KdTreeGpu):
        return "kd"
    if isinstance(struct, BallTreeGpu):
        return "ball"
    if isinstance(struct, BruteForceGpu):
        return "bf16" if form == FORM_BF16 else "brute"
    raise TypeError(f"not a knn structure: {type(struct)!r}")


def table_key(cls: str, dims: int, k: int, n: int) -> str:
    return f"{cls}:D{dims}:K{k}:N2^{max(0, math.ceil(math.log2(max(1, n))))}"


def candidates(cls: str, dims: int, k: int) -> list[dict]:
    """Result-neutral option sets for a class; the first is the built-in default."""
    if cls == "brute":
        return [dict(BLOCK_SIZE=b) for b in (256, 64, 128, 512, 1024)]
    if cls == "bf16":
        return [dict(tile=t, form=FORM_BF16) for t in (0, 1, 2, 4)]
    return [dict(BLOCK_SIZE=b, leaf_span_delta=s) for b, s in itertools.product((64, 128, 256), (0, -1, 1, 2, 3))]


def _default_span(cls: str, info: dict, k: int) -> int:
    """The library's default leaf span of the wave-cooperative traversal
    (knn_internal.hpp leaf_span_default)."""
    target = k
    if cls == "ball" and info["dims"] >= 8 and k >= 2:
        target = k // 2
    if cls == "kd":
        target = max(2 * k, 32)
    mc = max(1, info["max_children"])
    s = 0
    while target > (mc << s):
        s += 1
    return min(s, info["node_depth"])


def _resolve(cls: str, struct, k: int, cand: dict) -> dict:
    """Candidate -> explicit batchKnn keyword arguments (no field left for the
    table to fill: the leaf span made absolute, the tile given)."""
    opts = {key: v for key, v in cand.items() if key != "leaf_span_delta"}
    if cls in ("kd", "ball"):
        info = struct.info()
        span = _default_span(cls, info, k) + cand.get("leaf_span_delta", 0)
        span = max(0, min(span, info["node_depth"]))
        opts["leaf_span"] = 0xFFFFFFFF if span == 0 else span
    if cls == "bf16" and not opts.get("tile"):
        opts["tile"] = 8  # the widest staging (clamped to what fits)
    if "BLOCK_SIZE" not in opts:
        opts["BLOCK_SIZE"] = BUILTIN_BLOCK[cls]
    return opts


def tune(struct, queries: Any, k: int, max_distance2: float = float("inf"), form: int = FORM_PLAIN,
         base: dict | None = None, cands: Iterable[dict] | None = None, repeats: int = 3) -> dict:
    """Time every candidate on `queries` (a device tensor [q][D]) with the
    caller's semantic options `base` held fixed and return {"best": options,
    "ms": best time, "default_ms": time of the built-in defaults, "results":
    [(options, ms), ...]}.  Raises KnnError if two candidates return
    different rows."""
    import torch

    cls = class_of(struct, form)
    dims = int(queries.shape[1])
    nq = int(queries.shape[0])
    cands = list(cands) if cands is not None else candidates(cls, dims, k)
    base = dict(base or {})
    dev = queries.device
    ri = torch.empty((nq, k), dtype=torch.int32, device=dev)
    rd = torch.empty((nq, k), dtype=torch.float32, device=dev)
    q = Query(queries, nq, max_distance2, ri, rd)
    ref = None
    results = []
    # Short calls on a fresh structure run slower for their first few dozen
    # repetitions, which would penalise whichever candidate is timed first:
    # warm up with the built-in defaults for ~20 ms (at least 3 calls).
    t_end = time.perf_counter() + 0.02
    try:
        for i in range(200):
            struct.batchKnn(q, k, **{**base, **_resolve(cls, struct, k, cands[0])})
            torch.cuda.synchronize()
            if i >= 2 and time.perf_counter() > t_end:
                break
    except KnnError:
        pass  # the candidate loop below reports or skips it
    for cand in cands:
        opts = {**base, **_resolve(cls, struct, k, cand)}
        try:
            struct.batchKnn(q, k, **opts)
        except KnnError as e:
            if e.status in (_lib.KNN_ERR_UNSUPPORTED, _lib.KNN_ERR_INVALID_ARG):
                continue  # e.g. an LDS request above the workgroup limit
            raise
        torch.cuda.synchronize()
        got = (ri.clone(), rd.clone())
        if ref is None:
            ref = got
        elif not (torch.equal(ref[0], got[0]) and torch.equal(ref[1].view(torch.int32), got[1].view(torch.int32))):
            raise KnnError(_lib.KNN_ERR_INTERNAL, f"option set {cand} changed the result")
        times = []
        for _ in range(repeats):
            struct.batchKnn(q, k, **opts)
            times.append(struct.last_query_stats()["query_ms"])
        results.append((cand, min(times)))
    if not results:
        raise KnnError(_lib.KNN_ERR_UNSUPPORTED, "no candidate option set ran")
    best = min(results, key=lambda r: r[1])
    return {"best": best[0], "ms": best[1], "default_ms": results[0][1], "results": results}


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


def tuned_options(struct, k: int, form: int = FORM_PLAIN, table: dict | None = None,
                  path: str = DEFAULT_TABLE) -> dict:
    """Explicit batchKnn keyword arguments for `struct` from a table ({} when
    the bucket was never tuned)."""
    cls = class_of(struct, form)
    table = load_table(path) if table is None else table
    entry = table.get(table_key(cls, struct.dims, k, struct.size))
    return _resolve(cls, struct, k, entry["best"]) if entry else {}


def library_lookup(cls: str, dims: int, k: int, n: int) -> dict | None:
    """The entry compiled into the library for (class, D, K, n), or None."""
    out = (ctypes.c_int32 * 4)()
    _lib.check(_lib.load().knn_tuned_lookup(CLASSES[cls], dims, k, n, out))
    if not out[0]:
        return None
    return {"BLOCK_SIZE": out[1], "leaf_span_delta": out[2], "tile": out[3]}


def tune_and_store(struct, queries: Any, k: int, max_distance2: float = float("inf"), form: int = FORM_PLAIN,
                   path: str = DEFAULT_TABLE, **kw) -> dict:
    """Tune and record the winner under (class, D, K, size bucket)."""
    r = tune(struct, queries, k, max_distance2, form=form, **kw)
    table = load_table(path)
    table[table_key(class_of(struct, form), struct.dims, k, struct.size)] = {
        "best": r["best"], "ms": round(r["ms"], 4), "default_ms": round(r["default_ms"], 4),
        "queries": int(queries.shape[0]), "candidates": len(r["results"])}
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


def sweep_shapes() -> list[dict]:
    """The shapes the library is benchmarked on: the batched workload's
    (D, K, size) grid (uniform points, queries = size / 4) and the
    single-structure workloads at their own sizes and data."""
    shapes = []
    for b in range(10, 21):
        for k in (4, 8, 16):
            shapes.append(dict(cls="kd", dims=3, k=k, n=1 << b, nq=(1 << b) // 4, data="uniform", seed=4))
            shapes.append(dict(cls="bf16", dims=16, k=k, n=1 << b, nq=(1 << b) // 4, data="uniform", seed=4))
    shapes += [
        dict(cls="brute", dims=3, k=8, n=65536, nq=16384, data="uniform", seed=0, heap=True),
        dict(cls="kd", dims=3, k=8, n=65536, nq=16384, data="uniform", seed=0, heap=True),
        dict(cls="kd", dims=3, k=16, n=1 << 24, nq=1 << 22, data="gaussian3", seed=1),
        dict(cls="ball", dims=3, k=16, n=1 << 24, nq=1 << 22, data="gaussian3", seed=1),
        dict(cls="kd", dims=8, k=32, n=1 << 22, nq=1 << 20, data="aniso8", seed=3, repeats=2),
        dict(cls="ball", dims=8, k=32, n=1 << 22, nq=1 << 20, data="aniso8", seed=3, repeats=2),
    ]
    for d in (32, 64, 128):
        shapes.append(dict(cls="bf16", dims=d, k=32, n=1 << 20, nq=1 << 16, data="scaled", seed=2, repeats=2))
    return shapes


def _data(shape: dict):
    from . import datasets
    n, nq, d, seed = shape["n"], shape["nq"], shape["dims"], shape["seed"]
    if shape["data"] == "gaussian3":
        return datasets.gaussian_mixture_3d(n, nq, seed=seed)
    if shape["data"] == "aniso8":
        return datasets.anisotropic_mixture_8d(n, nq, seed=seed)
    if shape["data"] == "scaled":
        return datasets.scaled_normal(n, nq, d, seed=seed)
    return datasets.uniform_cube(n, nq, d, seed=seed)


def sweep(path: str = DEFAULT_TABLE, only: str | None = None) -> dict:
    """Tune every sweep shape and store the winners in `path`."""
    import torch

    dev = torch.device("cuda:0")
    table = load_table(path)
    for shape in sweep_shapes():
        if only and shape["cls"] != only:
            continue
        t0 = time.time()
        pts, qs = _data(shape)
        p = torch.from_numpy(pts).to(dev)
        q = torch.from_numpy(qs).to(dev)
        del pts, qs
        cls = shape["cls"]
        s = KdTreeGpu(p, depth_skips=0) if cls == "kd" else BallTreeGpu(p, depth_skips=1) if cls == "ball" \
            else BruteForceGpu(p)
        base = {"USE_HEAP_TRACKING": True} if shape.get("heap") else {}
        form = FORM_BF16 if cls == "bf16" else FORM_PLAIN
        r = tune(s, q, shape["k"], form=form, base=base, repeats=shape.get("repeats", 3))
        key = table_key(cls, shape["dims"], shape["k"], shape["n"])
        table[key] = {"best": r["best"], "ms": round(r["ms"], 4), "default_ms": round(r["default_ms"], 4),
                      "queries": shape["nq"], "candidates": len(r["results"])}
        save_table(table, path)
        print(f"[tune] {key}: best {r['best']} {r['ms']:.4f} ms (built-in default {r['default_ms']:.4f} ms), "
              f"{time.time() - t0:.1f} s", flush=True)
        s.close()
        del p, q
        torch.cuda.empty_cache()
    return table


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--sweep", action="store_true", help="re-measure the table on the benchmark shapes")
    ap.add_argument("--only", default=None, help="restrict the sweep to one class (brute, bf16, kd, ball)")
    ap.add_argument("--out", default=DEFAULT_TABLE)
    a = ap.parse_args(argv)
    if not a.sweep:
        ap.print_help()
        return 0
    sweep(a.out, a.only)
    return 0


if __name__ == "__main__":
    sys.exit(main())
