"""cudaexactknn_amd -- exact kNN on AMD MI355X (gfx950).

Python mirror of the reference's KnnTrees GPU API (README.md:52-114 of
rezvoj/CudaExactKnn), driving hand-written HIP kernels through the C ABI of
libknn_mi355x.so.  PyTorch is used only to hold device memory: points,
queries and results are CUDA/HIP tensors (or raw device pointers).

    bf = BruteForceGpu(d_points)                       # BruteForceGpu<D>(points, n)
    kd = KdTreeGpu(d_points, depth_skips=0, split_dimensions=D, block_size=256)
    ball = BallTreeGpu(d_points, depth_skips=1, block_size=256)
    q = Query(d_queries, count, max_distance2, r_indexes, r_distances)   # Query<D,K>
    kd.batchKnn(q, K, BLOCK_SIZE=256, USE_HEAP_TRACKING=False,
                USE_TRACKING_SHMEM=False, USE_STACK_SHMEM=False)

Semantics follow the reference's CPU BruteForce bit for bit (see
include/knn_mi355x.h): indices and squared distances are identical for every
structure and option; heap mode emits keys in descending order.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Any

from . import _lib
from ._lib import KnnError, QueryOpts, TreeInfo  # noqa: F401

__all__ = [
    "BruteForceGpu", "KdTreeGpu", "BallTreeGpu", "Query", "KnnError", "pinned_malloc", "pinned_free",
    "FORM_PLAIN", "FORM_BF16", "FORM_LEARNED", "library_path", "version", "Batch", "BatchItem",
    "KIND_BRUTE", "KIND_KD", "KIND_BALL", "pinned_array",
]

FORM_PLAIN, FORM_BF16, FORM_LEARNED = _lib.KNN_FORM_PLAIN, _lib.KNN_FORM_BF16, _lib.KNN_FORM_LEARNED


def library_path() -> str:
    return _lib.LIB_PATH


def version() -> int:
    return int(_lib.load().knn_version())


def _ptr(x: Any, dtype: Any = None, what: str = "tensor", device: Any = None) -> int:
    """Device pointer of a torch tensor, or an int passed through unchecked.
    A tensor must live on the GPU, be contiguous, have `dtype` when one is
    given (torch.float32 for points, queries and distances, torch.int32 for
    indices) and sit on `device` when one is given."""
    if x is None:
        return 0
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        _dtype(x, dtype, what)
        if not x.is_cuda:
            raise KnnError(_lib.KNN_ERR_INVALID_ARG, f"{what} must live on the GPU")
        if not x.is_contiguous():
            raise KnnError(_lib.KNN_ERR_INVALID_ARG, f"{what} must be contiguous")
        if device is not None and x.device != device:
            raise KnnError(_lib.KNN_ERR_INVALID_ARG, f"{what} is on {x.device}, the structure on {device}")
        return int(x.data_ptr())
    raise TypeError(f"expected a device tensor or pointer, got {type(x)!r}")


def _dtype(x: Any, dtype: Any, what: str) -> None:
    """A tensor (not a raw pointer) must have torch.<dtype>."""
    if dtype is not None and hasattr(x, "dtype") and hasattr(x, "data_ptr"):
        import torch
        if x.dtype != getattr(torch, dtype):
            raise KnnError(_lib.KNN_ERR_INVALID_ARG, f"{what} must be torch.{dtype}, got {x.dtype}")


def _rows(x: Any, count: int, what: str) -> None:
    """A tensor (not a raw pointer) must hold at least `count` rows."""
    if hasattr(x, "shape") and int(x.shape[0]) < count:
        raise KnnError(_lib.KNN_ERR_INVALID_ARG, f"{what} has {int(x.shape[0])} rows, {count} are read")


def _result_size(x: Any, q: int, K: int, what: str) -> None:
    """A result tensor must hold q*K elements of 4 bytes."""
    if hasattr(x, "numel") and x.numel() * x.element_size() < 4 * q * K:
        raise KnnError(_lib.KNN_ERR_INVALID_ARG, f"{what} holds fewer than count*K 4-byte elements")


def _shape2(x: Any, what: str):
    if not hasattr(x, "shape") or len(x.shape) != 2:
        raise KnnError(_lib.KNN_ERR_INVALID_ARG, f"{what} must be a 2-D [count, dims] tensor")
    return int(x.shape[0]), int(x.shape[1])


@dataclass
class Query:
    """Query<D, K> = {points, count, maxDistance2, rIndexes, rDistances}
    (src/KnnTrees.cuh:224-231): `points` [count, D] float32, `max_distance2` a
    SQUARED radius, results [count, K] int32 / float32.  `count` may be None
    (= rows of `points`)."""
    points: Any
    count: int | None
    max_distance2: float
    r_indexes: Any
    r_distances: Any

    def n(self) -> int:
        if self.count is not None:
            return int(self.count)
        return int(self.points.shape[0])


class _Structure:
    _kind = None

    def __init__(self):
        self._h = ctypes.c_void_p(0)

    @property
    def handle(self) -> int:
        return self._h.value or 0

    def batchKnnHost(self, queries: Any, K: int, max_distance2: float = float("inf"), r_indexes: Any = None,
                     r_distances: Any = None, chunk: int = 0, **opts) -> tuple:
        """Queries in host memory (numpy float32 [q][D]); results returned in
        (or written to) host int32 / float32 [q][K] arrays.  Uploads, searches
        and downloads overlap across chunks (knn_query_host).  Arrays in
        pinned memory (pinned_array) let the copies run concurrently."""
        import numpy as np
        qs = np.ascontiguousarray(queries, dtype=np.float32)
        if qs.ndim != 2 or qs.shape[1] != self.dims:
            raise KnnError(_lib.KNN_ERR_INVALID_ARG, f"queries must be [q][{self.dims}] float32")
        q = qs.shape[0]
        ri = np.empty((q, K), np.int32) if r_indexes is None else r_indexes
        rd = np.empty((q, K), np.float32) if r_distances is None else r_distances
        for a, dt in ((ri, np.int32), (rd, np.float32)):
            if a.dtype != dt or a.shape != (q, K) or not a.flags.c_contiguous:
                raise KnnError(_lib.KNN_ERR_INVALID_ARG, "result arrays must be C-contiguous [q][K]")
        o = _opts(opts.pop("BLOCK_SIZE", 0), opts.pop("USE_HEAP_TRACKING", False), opts.pop("USE_TRACKING_SHMEM", False),
                  opts.pop("USE_STACK_SHMEM", False), **opts)
        _lib.check(_lib.load().knn_query_host(self._h, qs.ctypes.data, q, float(max_distance2), ri.ctypes.data,
                                              rd.ctypes.data, int(K), ctypes.byref(o), int(chunk)))
        return ri, rd

    def reserve(self, q: int, K: int, **opts) -> None:
        """Size the structure's query scratch for any batch of up to q
        queries with this K and these options (knn_query_reserve), so that a
        later batchKnn(..., stream=s) allocates nothing and can be captured
        into a HIP graph.  Options as for batchKnn; stream is ignored."""
        o = _opts(opts.pop("BLOCK_SIZE", 0), opts.pop("USE_HEAP_TRACKING", False), opts.pop("USE_TRACKING_SHMEM", False),
                  opts.pop("USE_STACK_SHMEM", False), **opts)
        _lib.check(_lib.load().knn_query_reserve(self._h, int(q), int(K), ctypes.byref(o)))

    def info(self) -> dict:
        ti = TreeInfo()
        _lib.check(_lib.load().knn_tree_get_info(self._h, ctypes.byref(ti)))
        return ti.as_dict()

    def _query(self, query: Query, K: int, opts: QueryOpts) -> None:
        q = query.n()
        if hasattr(query.points, "shape") and q:
            rows, d = _shape2(query.points, "query points")
            if d != self.dims:
                raise KnnError(_lib.KNN_ERR_INVALID_ARG, f"query dims {d} != structure dims {self.dims}")
            _rows(query.points, q, "query points")
        if q < 0 or K < 1:
            raise KnnError(_lib.KNN_ERR_INVALID_ARG, "count must be >= 0 and K >= 1")
        _dtype(query.points, "float32", "query points")
        _dtype(query.r_indexes, "int32", "r_indexes")
        _dtype(query.r_distances, "float32", "r_distances")
        _result_size(query.r_indexes, q, K, "r_indexes")
        _result_size(query.r_distances, q, K, "r_distances")
        dv = getattr(self, "device", None)
        _lib.check(_lib.load().knn_query(self._h, _ptr(query.points, "float32", "query points", dv), q,
                                         float(query.max_distance2),
                                         _ptr(query.r_indexes, "int32", "r_indexes", dv),
                                         _ptr(query.r_distances, "float32", "r_distances", dv), int(K),
                                         ctypes.byref(opts)))

    def last_query_stats(self) -> dict:
        buf = (ctypes.c_float * 8)()
        _lib.check(_lib.load().knn_last_query_stats(self._h, buf))
        return {"rerun_queries": buf[0], "bf16_survivor_fraction": buf[1], "radius_ratio": buf[2],
                "query_ms": buf[3], "kernel_ms": buf[4], "node_bytes": buf[5], "leaf_slots": buf[6],
                "order_ms": buf[7], "bf16_seeded": buf[5], "far_fraction": buf[1]}

    def close(self) -> None:
        if self._h and self._h.value:
            _lib.load().knn_destroy(self._h)
            self._h = ctypes.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _opts(BLOCK_SIZE, heap, track, stack, form=FORM_PLAIN, sort_queries=True, leaf_span=0, stream=None,
          count_visits=False, per_lane=False, morton=False, xcd_remap=True, tile=0, learned_stats=False,
          learned_radius=False):
    o = QueryOpts()
    o.block = int(BLOCK_SIZE or 0)
    o.heap = int(bool(heap))
    o.track_lds = int(bool(track))
    o.stack_lds = int(bool(stack))
    o.form = int(form)
    o.flags = ((0 if sort_queries else _lib.KNN_FLAG_NO_QUERY_SORT) | (_lib.KNN_FLAG_COUNT_VISITS if count_visits else 0)
               | (_lib.KNN_FLAG_PER_LANE if per_lane else 0) | (_lib.KNN_FLAG_MORTON_ORDER if morton else 0)
               | (0 if xcd_remap else _lib.KNN_FLAG_NO_XCD_REMAP)
               | (_lib.KNN_FLAG_LEARNED_STATS if learned_stats else 0)
               | (_lib.KNN_FLAG_LEARNED_RADIUS if learned_radius else 0))
    o.leaf_span = int(leaf_span)
    o.tile = int(tile)
    o.stream = stream
    return o


class BruteForceGpu(_Structure):
    """BruteForceGpu<D>(const Array<float,D>* points, uint size)
    (src/BruteForceGpu.cuh:64-68): copies the device points."""

    def __init__(self, points: Any, size: int | None = None):
        super().__init__()
        n, d = _shape2(points, "points")
        n = n if size is None else int(size)
        _rows(points, n, "points")
        self.dims, self.size, self.device = d, n, getattr(points, "device", None)
        _lib.check(_lib.load().knn_bf_create(_ptr(points, "float32", "points"), n, d, ctypes.byref(self._h)))

    def batchKnn(self, query: Query, K: int, BLOCK_SIZE: int = 0, USE_HEAP_TRACKING: bool = False,
                 USE_TRACKING_SHMEM: bool = False, form: int = FORM_PLAIN, stream=None, tile: int = 0) -> None:
        """batchKnn<K, BLOCK_SIZE, USE_HEAP_TRACKING, USE_TRACKING_SHMEM>(Query&)
        (src/BruteForceGpu.cuh:74-83).  `tile`: bf16 form, point tiles staged per barrier."""
        self._query(query, K, _opts(BLOCK_SIZE, USE_HEAP_TRACKING, USE_TRACKING_SHMEM, False, form,
                                    stream=stream, tile=tile))


class _Tree(_Structure):
    def batchKnn(self, query: Query, K: int, BLOCK_SIZE: int = 0, USE_HEAP_TRACKING: bool = False,
                 USE_TRACKING_SHMEM: bool = False, USE_STACK_SHMEM: bool = False, form: int = FORM_PLAIN,
                 sort_queries: bool = True, leaf_span: int = 0, stream=None, count_visits: bool = False,
                 per_lane: bool = False, morton: bool = False, xcd_remap: bool = True,
                 learned_stats: bool = False, learned_radius: bool = False) -> None:
        """batchKnn<K, BLOCK_SIZE, USE_HEAP_TRACKING, USE_TRACKING_SHMEM, USE_STACK_SHMEM>(Query&)
        (src/KdTreeGpu.cuh:393-410, src/BallTreeGpu.cuh:631-648).  Learned form
        (form=FORM_LEARNED): `learned_radius` also starts each traversal from
        the predicted radius."""
        self._query(query, K, _opts(BLOCK_SIZE, USE_HEAP_TRACKING, USE_TRACKING_SHMEM, USE_STACK_SHMEM, form,
                                    sort_queries, leaf_span, stream, count_visits, per_lane, morton, xcd_remap,
                                    learned_stats=learned_stats, learned_radius=learned_radius))

    def build_ms(self) -> float:
        return float(_lib.load().knn_last_build_ms(self._h))

    def export(self):
        """Copy the tree to host numpy arrays: (nodes, leaf_index, leaf_points, dims).
        `dims` holds each node's split dimension for ball trees and zeros for kd
        trees (whose split dimension is the level modulo SPLIT_DIMENSIONS)."""
        import numpy as np
        i = self.info()
        nodes = np.empty(i["node_size"] * i["node_bytes"] // 4, np.float32)
        leaves = np.empty(i["index_size"] * (1 + i["dims"]), np.float32)
        dims = np.zeros(max(1, i["node_size"]), np.uint32)
        _lib.check(_lib.load().knn_tree_export(self._h, nodes.ctypes.data, leaves.ctypes.data, dims.ctypes.data))
        leaves = leaves.reshape(i["index_size"], 1 + i["dims"])
        leaf_index = leaves[:, 0].copy().view(np.int32)
        leaf_points = leaves[:, 1:].copy()
        return nodes, leaf_index, leaf_points, dims[: i["node_size"]]

    def serialize(self) -> bytes:
        """The tree in the library's stable serialized format (knn_tree_serialize)."""
        L = _lib.load()
        n = ctypes.c_size_t(0)
        _lib.check(L.knn_tree_serialized_size(self._h, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _lib.check(L.knn_tree_serialize(self._h, buf, n.value))
        return buf.raw

    def save(self, path: str) -> None:
        with open(path, "wb") as f:
            f.write(self.serialize())

    @staticmethod
    def from_bytes(data: bytes) -> "_Tree":
        """A KdTreeGpu or BallTreeGpu restored from serialize() output."""
        h = ctypes.c_void_p(0)
        buf = ctypes.create_string_buffer(bytes(data), len(data))
        _lib.check(_lib.load().knn_tree_deserialize(buf, len(data), ctypes.byref(h)))
        ti = TreeInfo()
        _lib.check(_lib.load().knn_tree_get_info(h, ctypes.byref(ti)))
        cls = KdTreeGpu if ti.kind == _lib.KNN_KIND_KD else BallTreeGpu
        t = cls.__new__(cls)
        _Structure.__init__(t)
        t._h = h
        t.dims, t.size = ti.dims, ti.n
        return t

    @staticmethod
    def load(path: str) -> "_Tree":
        with open(path, "rb") as f:
            return _Tree.from_bytes(f.read())

    def train_radius(self, sample: Any, K: int, epochs: int = 0, lr: float = 0.0) -> dict:
        """Train the learned-radius predictor (KNN_FORM_LEARNED) on sample queries."""
        n, d = _shape2(sample, "sample")
        stats = (ctypes.c_float * 2)()
        if d != self.dims:
            raise KnnError(_lib.KNN_ERR_INVALID_ARG, f"sample dims {d} != structure dims {self.dims}")
        _lib.check(_lib.load().knn_learned_train(self._h, _ptr(sample, "float32", "sample", getattr(self, "device", None)),
                                                 n, int(K), int(epochs), float(lr), stats))
        return {"loss": stats[0], "train_ms": stats[1]}


class KdTreeGpu(_Tree):
    """KdTreeGpu<D, SPLIT_DIMENSIONS, BLOCK_SIZE>(points, size, depthSkips = 0)
    (src/KdTreeGpu.cuh:375-385)."""

    def __init__(self, points: Any, size: int | None = None, depth_skips: int = 0,
                 split_dimensions: int | None = None, block_size: int = 256):
        super().__init__()
        n, d = _shape2(points, "points")
        n = n if size is None else int(size)
        sd = d if split_dimensions is None else int(split_dimensions)
        _rows(points, n, "points")
        self.dims, self.size, self.device = d, n, getattr(points, "device", None)
        _lib.check(_lib.load().knn_kd_build(_ptr(points, "float32", "points"), n, d, sd, int(depth_skips),
                                            int(block_size),
                                            ctypes.byref(self._h)))


class BallTreeGpu(_Tree):
    """BallTreeGpu<D, BLOCK_SIZE>(points, size, depthSkips = 1)
    (src/BallTreeGpu.cuh:614-624)."""

    def __init__(self, points: Any, size: int | None = None, depth_skips: int = 1, block_size: int = 256):
        super().__init__()
        n, d = _shape2(points, "points")
        n = n if size is None else int(size)
        _rows(points, n, "points")
        self.dims, self.size, self.device = d, n, getattr(points, "device", None)
        _lib.check(_lib.load().knn_ball_build(_ptr(points, "float32", "points"), n, d, int(depth_skips),
                                              int(block_size),
                                              ctypes.byref(self._h)))


def pinned_malloc(nbytes: int) -> int:
    """pinnedMalloc (src/KnnTrees.cuh:233-238): page-locked host memory."""
    p = ctypes.c_void_p(0)
    _lib.check(_lib.load().knn_pinned_alloc(int(nbytes), ctypes.byref(p)))
    return p.value or 0


def pinned_free(ptr: int) -> None:
    """pinnedFree (src/KnnTrees.cuh:240-243)."""
    _lib.check(_lib.load().knn_pinned_free(ctypes.c_void_p(ptr)))


class pinned_array:
    """A numpy array in page-locked host memory (knn_pinned_alloc), freed by
    close() or on garbage collection."""

    def __init__(self, shape, dtype):
        import numpy as np
        self.dtype = np.dtype(dtype)
        n = int(np.prod(shape)) * self.dtype.itemsize
        self.ptr = pinned_malloc(max(1, n))
        buf = (ctypes.c_uint8 * max(1, n)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=self.dtype, count=int(np.prod(shape))).reshape(shape)

    def close(self) -> None:
        if self.ptr:
            self.array = None
            pinned_free(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


KIND_BRUTE, KIND_KD, KIND_BALL = _lib.KNN_KIND_BRUTE, _lib.KNN_KIND_KD, _lib.KNN_KIND_BALL


@dataclass
class BatchItem:
    """One dataset of a Batch: `kind` KIND_KD / KIND_BALL (dims 2, 3, 4, 8) or
    KIND_BRUTE (bf16 form, dims <= 256), K <= 32.  Tensors stay referenced
    by the item: every replay reads points and queries and writes results."""
    kind: int
    points: Any
    queries: Any
    k: int
    r_indexes: Any
    r_distances: Any
    max_distance2: float = float("inf")
    split_dimensions: int = 0
    depth_skips: int = 0


class Batch:
    """Many independent datasets built and queried by two captured HIP graphs
    (knn_batch_* in include/knn_mi355x.h); the queries run as one persistent
    launch over (dataset, query block) work items."""

    STAT_NAMES = ("build_ms", "query_ms", "overflow_queries", "work_items", "workgroups", "build_graph_nodes",
                  "query_graph_nodes", "queries")

    def __init__(self, items: list, build_streams: int = 0, blocks_per_cu: int = 0, tile: int = 0,
                 leaf_span: int = 0):
        self._h = ctypes.c_void_p(0)
        self.items = list(items)
        arr = (_lib.BatchItem * max(1, len(self.items)))()
        for j, it in enumerate(self.items):
            n, d = _shape2(it.points, "points")
            nq, dq = _shape2(it.queries, "queries")
            if dq != d:
                raise KnnError(_lib.KNN_ERR_INVALID_ARG, f"batch item {j}: queries have {dq} dims, points {d}")
            a = arr[j]
            a.kind, a.dims, a.n, a.nq, a.k = int(it.kind), d, n, nq, int(it.k)
            a.split_dims, a.depth_skips, a.max_d2 = int(it.split_dimensions), int(it.depth_skips), float(it.max_distance2)
            dv = getattr(it.points, "device", None)
            for t, dt, nm in ((it.points, "float32", "points"), (it.queries, "float32", "queries"),
                              (it.r_indexes, "int32", "r_indexes"), (it.r_distances, "float32", "r_distances")):
                _dtype(t, dt, f"batch item {j}: {nm}")
            _result_size(it.r_indexes, nq, int(it.k), f"batch item {j}: r_indexes")
            _result_size(it.r_distances, nq, int(it.k), f"batch item {j}: r_distances")
            a.d_pts = _ptr(it.points, "float32", f"batch item {j}: points", dv)
            a.d_q = _ptr(it.queries, "float32", f"batch item {j}: queries", dv)
            a.d_idx = _ptr(it.r_indexes, "int32", f"batch item {j}: r_indexes", dv)
            a.d_dist = _ptr(it.r_distances, "float32", f"batch item {j}: r_distances", dv)
        o = _lib.BatchOpts(int(build_streams), int(blocks_per_cu), int(tile), int(leaf_span))
        _lib.check(_lib.load().knn_batch_create(arr, len(self.items), ctypes.byref(o), ctypes.byref(self._h)))

    def run(self, build: bool = True, query: bool = True, stream=None) -> None:
        """Replay the build graph, the query graph, or both (synchronous when
        stream is None)."""
        what = (1 if build else 0) | (2 if query else 0)
        _lib.check(_lib.load().knn_batch_run(self._h, what, stream))

    def stats(self) -> dict:
        out = (ctypes.c_float * 8)()
        _lib.check(_lib.load().knn_batch_stats(self._h, out))
        return dict(zip(self.STAT_NAMES, list(out)))

    def close(self) -> None:
        if self._h:
            _lib.check(_lib.load().knn_batch_destroy(self._h))
            self._h = ctypes.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
