"""ctypes binding of libknn_mi355x.so (the C ABI in include/knn_mi355x.h).

The library is built in-tree by cudaexactknn_amd.build; there is no CPU
fallback: if the shared object is missing, loading fails loudly.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libknn_mi355x.so")

KNN_OK = 0
KNN_ERR_INVALID_ARG = 1
KNN_ERR_HIP = 2
KNN_ERR_OOM = 3
KNN_ERR_UNSUPPORTED = 4
KNN_ERR_INTERNAL = 5

KNN_KIND_BRUTE, KNN_KIND_KD, KNN_KIND_BALL = 0, 1, 2
KNN_FORM_PLAIN, KNN_FORM_BF16, KNN_FORM_LEARNED = 0, 1, 2
KNN_FLAG_NO_QUERY_SORT = 0x1
KNN_FLAG_COUNT_VISITS = 0x2
KNN_FLAG_PER_LANE = 0x4
KNN_FLAG_MORTON_ORDER = 0x8
KNN_FLAG_NO_XCD_REMAP = 0x10
KNN_FLAG_LEARNED_STATS = 0x20
KNN_FLAG_LEARNED_RADIUS = 0x40

# Every symbol include/knn_mi355x.h declares.
EXPORTED = (
    "knn_version", "knn_bf_create", "knn_kd_build", "knn_ball_build", "knn_query", "knn_destroy",
    "knn_last_error", "knn_pinned_alloc", "knn_pinned_free", "knn_tree_get_info", "knn_tree_export",
    "knn_last_build_ms", "knn_learned_train", "knn_last_query_stats", "knn_mlp_param_count", "knn_mlp_forward",
    "knn_mlp_gradient", "knn_batch_create", "knn_batch_run", "knn_batch_stats", "knn_batch_destroy",
    "knn_tree_serialized_size", "knn_tree_serialize", "knn_tree_deserialize", "knn_query_host",
    "knn_query_noalloc", "knn_query_reserve", "knn_query_resolved_opts", "knn_tuned_lookup",
)


class QueryOpts(ctypes.Structure):
    _fields_ = [
        ("block", ctypes.c_uint32), ("heap", ctypes.c_uint32), ("track_lds", ctypes.c_uint32),
        ("stack_lds", ctypes.c_uint32), ("form", ctypes.c_uint32), ("flags", ctypes.c_uint32),
        ("leaf_span", ctypes.c_uint32), ("tile", ctypes.c_uint32), ("stream", ctypes.c_void_p),
    ]


class BatchItem(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_uint32), ("dims", ctypes.c_uint32), ("n", ctypes.c_uint32), ("nq", ctypes.c_uint32),
        ("k", ctypes.c_uint32), ("split_dims", ctypes.c_uint32), ("depth_skips", ctypes.c_uint32),
        ("max_d2", ctypes.c_float), ("d_pts", ctypes.c_void_p), ("d_q", ctypes.c_void_p),
        ("d_idx", ctypes.c_void_p), ("d_dist", ctypes.c_void_p),
    ]


class BatchOpts(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in ("build_streams", "blocks_per_cu", "tile", "leaf_span")]


class TreeInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in (
        "kind", "n", "dims", "split_dims", "depth_skips", "node_depth", "max_children", "node_size",
        "index_size", "node_bytes", "leaf_bytes", "reserved")]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class KnnError(RuntimeError):
    """Raised for a non-zero knn_status (the reference's CudaException,
    src/KnnTrees.cuh:20-23, is a std::runtime_error as well)."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[knn status {status}] {message}")
        self.status = status


_LIB = None


def load() -> ctypes.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise OSError(
            f"{LIB_PATH} is missing: build it with `python -m cudaexactknn_amd.build` "
            "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    P, u32, f32, st = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float, ctypes.c_int
    L.knn_version.restype = u32
    L.knn_version.argtypes = []
    L.knn_bf_create.argtypes = [P, u32, u32, ctypes.POINTER(P)]
    L.knn_kd_build.argtypes = [P, u32, u32, u32, u32, u32, ctypes.POINTER(P)]
    L.knn_ball_build.argtypes = [P, u32, u32, u32, u32, ctypes.POINTER(P)]
    L.knn_query.argtypes = [P, P, u32, f32, P, P, u32, ctypes.POINTER(QueryOpts)]
    L.knn_query_host.argtypes = [P, P, u32, f32, P, P, u32, ctypes.POINTER(QueryOpts), u32]
    L.knn_query_noalloc.argtypes = [P, P, u32, f32, P, P, u32, ctypes.POINTER(QueryOpts)]
    L.knn_query_reserve.argtypes = [P, u32, u32, ctypes.POINTER(QueryOpts)]
    L.knn_query_resolved_opts.argtypes = [P, u32, u32, ctypes.POINTER(QueryOpts), ctypes.POINTER(QueryOpts)]
    L.knn_tuned_lookup.argtypes = [u32, u32, u32, u32, ctypes.POINTER(ctypes.c_int32)]
    L.knn_destroy.argtypes = [P]
    L.knn_last_error.restype = ctypes.c_char_p
    L.knn_last_error.argtypes = []
    L.knn_pinned_alloc.argtypes = [ctypes.c_size_t, ctypes.POINTER(P)]
    L.knn_pinned_free.argtypes = [P]
    L.knn_tree_get_info.argtypes = [P, ctypes.POINTER(TreeInfo)]
    L.knn_tree_export.argtypes = [P, P, P, P]
    L.knn_last_build_ms.restype = f32
    L.knn_last_build_ms.argtypes = [P]
    L.knn_learned_train.argtypes = [P, P, u32, u32, u32, f32, P]
    L.knn_last_query_stats.argtypes = [P, P]
    L.knn_mlp_param_count.argtypes = [ctypes.POINTER(u32)]
    L.knn_mlp_forward.argtypes = [P, P, u32, u32, P]
    L.knn_mlp_gradient.argtypes = [P, P, P, u32, P, ctypes.POINTER(f32)]
    L.knn_batch_create.argtypes = [ctypes.POINTER(BatchItem), u32, ctypes.POINTER(BatchOpts), ctypes.POINTER(P)]
    L.knn_batch_run.argtypes = [P, u32, P]
    L.knn_batch_stats.argtypes = [P, P]
    L.knn_batch_destroy.argtypes = [P]
    L.knn_tree_serialized_size.argtypes = [P, ctypes.POINTER(ctypes.c_size_t)]
    L.knn_tree_serialize.argtypes = [P, P, ctypes.c_size_t]
    L.knn_tree_deserialize.argtypes = [P, ctypes.c_size_t, ctypes.POINTER(P)]
    for name in ("knn_bf_create", "knn_kd_build", "knn_ball_build", "knn_query", "knn_destroy",
                 "knn_pinned_alloc", "knn_pinned_free", "knn_tree_get_info", "knn_tree_export",
                 "knn_learned_train", "knn_last_query_stats", "knn_mlp_param_count", "knn_mlp_forward",
                 "knn_mlp_gradient", "knn_batch_create", "knn_batch_run", "knn_batch_stats", "knn_batch_destroy",
                 "knn_tree_serialized_size", "knn_tree_serialize", "knn_tree_deserialize", "knn_query_host",
                 "knn_query_noalloc", "knn_query_reserve", "knn_query_resolved_opts", "knn_tuned_lookup"):
        getattr(L, name).restype = st
    _LIB = L
    return L


def check(status: int) -> None:
    if status != KNN_OK:
        msg = load().knn_last_error()
        raise KnnError(status, msg.decode() if msg else "")
