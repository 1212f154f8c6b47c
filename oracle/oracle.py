"""ctypes front-end of the CPU oracle (liboracle_knn.so, built from knn_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  Every function
follows the reference file:line named in knn_oracle.c.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class TreeInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in (
        "n", "dims", "split_dims", "skips", "node_depth", "max_children", "node_size", "index_size")]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def build(force: bool = False) -> str:
    """Compile liboracle_knn.so with the committed Makefile."""
    so = os.path.join(_HERE, "liboracle_knn.so")
    if force or not os.path.exists(so) or os.path.getmtime(so) < os.path.getmtime(
            os.path.join(_HERE, "knn_oracle.c")):
        if force and os.path.exists(so):
            os.remove(so)
        subprocess.run(["make", "-s", "-C", _HERE, "liboracle_knn.so"], check=True)
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle_knn.so")
        if not os.path.exists(so):
            build()
        L = ctypes.CDLL(so)
        P = ctypes.c_void_p
        u32, i32, f32 = ctypes.c_uint32, ctypes.c_int, ctypes.c_float
        L.oracle_bf_knn.argtypes = [P, u32, u32, P, u32, f32, u32, i32, P, P, i32]
        L.oracle_kd_build.argtypes = [P, u32, u32, u32, u32]
        L.oracle_kd_build.restype = P
        L.oracle_kd_export.argtypes = [P, P, P]
        L.oracle_kd_knn.argtypes = [P, P, u32, f32, u32, i32, i32, P, P, P, i32]
        L.oracle_kd_free.argtypes = [P]
        L.oracle_ball_build.argtypes = [P, u32, u32, u32]
        L.oracle_ball_build.restype = P
        L.oracle_ball_export.argtypes = [P, P, P, P]
        L.oracle_ball_knn.argtypes = [P, P, u32, f32, u32, i32, i32, P, P, P, i32]
        L.oracle_ball_free.argtypes = [P]
        L.oracle_tree_get_info.argtypes = [P, ctypes.POINTER(TreeInfo)]
        L.oracle_dist2.argtypes = [P, P, u32]
        L.oracle_dist2.restype = f32
        L.oracle_ball_bound.argtypes = [P, P, f32, u32]
        L.oracle_ball_bound.restype = f32
        L.oracle_ordered_bits.argtypes = [f32]
        L.oracle_ordered_bits.restype = u32
        L.oracle_log2_ceil.argtypes = [u32]
        L.oracle_log2_ceil.restype = u32
        L.oracle_set_mutation.argtypes = [i32]
        L.oracle_set_raw_heap.argtypes = [i32]
        _LIB = L
    return _LIB


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def set_mutation(m: int) -> None:
    lib().oracle_set_mutation(int(m))


def set_raw_heap(on: bool) -> None:
    lib().oracle_set_raw_heap(1 if on else 0)


def bf_knn(points, queries, k, max_d2=np.inf, heap=False, threads=0):
    """BruteForce<D>::batchKnn<K, heap> (BruteForce.cuh:27-44)."""
    p, q = _f32(points), _f32(queries)
    n, d = p.shape
    nq = q.shape[0]
    idx = np.empty((nq, k), np.int32)
    dist = np.empty((nq, k), np.float32)
    rc = lib().oracle_bf_knn(_ptr(p), n, d, _ptr(q), nq, float(np.float32(max_d2)), k,
                             int(heap), _ptr(idx), _ptr(dist), threads)
    assert rc == 0
    return idx, dist


class _Tree:
    _free = None

    def info(self) -> dict:
        ti = TreeInfo()
        lib().oracle_tree_get_info(self._h, ctypes.byref(ti))
        return ti.as_dict()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            getattr(lib(), self._free)(h)
            self._h = None


class KdTree(_Tree):
    """KdTree<D>(points, n, depthSkips) with split dims `split_dims`
    (KdTree.cuh:27-71; the GPU variant's dim rule depth % SplitDims,
    KdTreeGpu.cuh:329)."""
    _free = "oracle_kd_free"

    def __init__(self, points, depth_skips=0, split_dims=None):
        p = _f32(points)
        self.n, self.d = p.shape
        sd = self.d if split_dims is None else split_dims
        self._h = lib().oracle_kd_build(_ptr(p), self.n, self.d, sd, depth_skips)
        assert self._h

    def export(self):
        i = self.info()
        nodes = np.empty(i["node_size"], np.float32)
        leaf = np.empty(i["index_size"], np.int32)
        lib().oracle_kd_export(self._h, _ptr(nodes), _ptr(leaf))
        return nodes, leaf

    def knn(self, queries, k, max_d2=np.inf, heap=False, strict_ref=False, threads=0, counters=False):
        q = _f32(queries)
        nq = q.shape[0]
        idx = np.empty((nq, k), np.int32)
        dist = np.empty((nq, k), np.float32)
        cnt = np.zeros(4, np.uint64)
        rc = lib().oracle_kd_knn(self._h, _ptr(q), nq, float(np.float32(max_d2)), k, int(heap),
                                 int(strict_ref), _ptr(idx), _ptr(dist), _ptr(cnt), threads)
        assert rc == 0
        return (idx, dist, cnt) if counters else (idx, dist)


class BallTree(_Tree):
    """BallTree<D>(points, n, depthSkips) (BallTree.cuh:27-137)."""
    _free = "oracle_ball_free"

    def __init__(self, points, depth_skips=1):
        p = _f32(points)
        self.n, self.d = p.shape
        self._h = lib().oracle_ball_build(_ptr(p), self.n, self.d, depth_skips)
        assert self._h

    def export(self):
        i = self.info()
        nodes = np.empty((i["node_size"], 2 * self.d + 2), np.float32)
        dims = np.empty(i["node_size"], np.uint32)
        leaf = np.empty(i["index_size"], np.int32)
        lib().oracle_ball_export(self._h, _ptr(nodes), _ptr(dims), _ptr(leaf))
        return nodes, dims, leaf

    def knn(self, queries, k, max_d2=np.inf, heap=False, strict_ref=False, threads=0, counters=False):
        q = _f32(queries)
        nq = q.shape[0]
        idx = np.empty((nq, k), np.int32)
        dist = np.empty((nq, k), np.float32)
        cnt = np.zeros(4, np.uint64)
        rc = lib().oracle_ball_knn(self._h, _ptr(q), nq, float(np.float32(max_d2)), k, int(heap),
                                   int(strict_ref), _ptr(idx), _ptr(dist), _ptr(cnt), threads)
        assert rc == 0
        return (idx, dist, cnt) if counters else (idx, dist)


def dist2(a, b) -> np.float32:
    a, b = _f32(a), _f32(b)
    return np.float32(lib().oracle_dist2(_ptr(a), _ptr(b), a.shape[0]))


def ball_bound(q, c, r, d) -> np.float32:
    q, c = _f32(q), _f32(c)
    return np.float32(lib().oracle_ball_bound(_ptr(q), _ptr(c), float(r), d))


def ordered_bits(f) -> int:
    return int(lib().oracle_ordered_bits(float(np.float32(f))))


def log2_ceil(x: int) -> int:
    return int(lib().oracle_log2_ceil(int(x)))
