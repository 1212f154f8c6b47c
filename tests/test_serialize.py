"""Serialized tree format: corrupted or inconsistent buffers are rejected
before any device work (CPU), and a round trip reproduces the tree and its
query results bit for bit (GPU)."""
import ctypes
import struct

import numpy as np
import pytest

import cudaexactknn_amd as knn
from cudaexactknn_amd import _lib


def _fnv1a(b: bytes) -> int:
    h = 0xcbf29ce484222325
    for x in b:
        h ^= x
        h = (h * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def _fake_kd(n=8, dims=3):
    """A header + zero payload that is self-consistent for an n-point kd tree."""
    # tree sizing (src/KdTree.cuh:56-62): L = max(1, ceil log2 n), skips 0
    L = max(1, (n - 1).bit_length())
    info = [_lib.KNN_KIND_KD, n, dims, dims, 0, L, 1, (1 << L) - 1, 1 << L, 4, 4 + 4 * dims, 0]
    node, dim, leaf, bbox = info[7] * 4, 0, info[8] * info[10], 2 * dims * 4
    payload = bytes(node + dim + leaf + bbox)
    hdr = b"KNNTREE1" + struct.pack("<II", 1, 0x01020304) + struct.pack("<12I", *info)
    hdr += struct.pack("<5Q", node, dim, leaf, bbox, _fnv1a(payload))
    assert len(hdr) == 104
    return bytearray(hdr + payload)


def _deserialize_status(buf) -> int:
    h = ctypes.c_void_p(0)
    raw = ctypes.create_string_buffer(bytes(buf), len(buf))
    return _lib.load().knn_tree_deserialize(raw, len(buf), ctypes.byref(h))


@pytest.mark.parametrize("corrupt", ["magic", "version", "truncated", "checksum", "info", "sizes", "short"])
def test_bad_buffers_rejected_without_device(corrupt):
    buf = _fake_kd()
    if corrupt == "magic":
        buf[0:8] = b"KNNTREE9"
    elif corrupt == "version":
        buf[8:12] = struct.pack("<I", 7)
    elif corrupt == "truncated":
        buf = buf[:-1]
    elif corrupt == "checksum":
        buf[-1] ^= 0x40
    elif corrupt == "info":
        buf[16 + 4 * 7:16 + 4 * 8] = struct.pack("<I", 99)  # node_size
    elif corrupt == "sizes":
        buf[64:72] = struct.pack("<Q", 12345)
    elif corrupt == "short":
        buf = buf[:50]
    want = _lib.KNN_ERR_UNSUPPORTED if corrupt == "version" else _lib.KNN_ERR_INVALID_ARG
    assert _deserialize_status(buf) == want


@pytest.mark.gpu
@pytest.mark.parametrize("kind,skips", [("kd", 0), ("kd", 2), ("ball", 1), ("ball", 3)])
def test_round_trip_bit_identical(oracle, cuda, tmp_path, kind, skips):
    import torch
    from tests.test_gpu_parity import dev, run, same
    g = np.random.Generator(np.random.Philox(91 + skips))
    pts = g.random((50000, 3), dtype=np.float32)
    qs = g.random((4000, 3), dtype=np.float32)
    p = dev(pts, cuda)
    t = knn.KdTreeGpu(p, depth_skips=skips) if kind == "kd" else knn.BallTreeGpu(p, depth_skips=skips)
    data = t.serialize()
    path = str(tmp_path / "tree.knn")
    t.save(path)
    t2 = knn.KdTreeGpu.load(path)
    assert type(t2) is type(t) and t2.info() == t.info()
    a, b = t.export(), t2.export()
    parts = 4 if kind == "ball" else 3  # kd trees have no split-dim array
    for x, y in zip(a[:parts], b[:parts]):
        assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))
    assert t2.serialize() == data
    want = oracle.bf_knn(pts, qs, 16)
    same(run(t2, qs, 16, cuda), want, f"{kind} restored from bytes")
    same(run(t, qs, 16, cuda), run(t2, qs, 16, cuda, per_lane=True), f"{kind} original vs restored")
    t.close()
    t2.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n,d,kind", [(1, 3, "kd"), (5, 7, "kd"), (1000, 3, "ball"), (300001, 3, "kd"),
                                      (70000, 5, "ball"), (20000, 40, "kd")])
def test_bounding_box_is_exact(cuda, n, d, kind):
    """The tree's bounding box (last 2 D floats of the serialized tree: minima,
    then maxima) equals the per-dimension extrema of the points."""
    from tests.test_gpu_parity import dev
    g = np.random.Generator(np.random.Philox(n + d))
    pts = (g.standard_normal((n, d)) * np.exp(g.uniform(-3, 3, d))).astype(np.float32)
    p = dev(pts, cuda)
    t = knn.KdTreeGpu(p, depth_skips=0) if kind == "kd" else knn.BallTreeGpu(p, depth_skips=1)
    bbox = np.frombuffer(t.serialize()[-8 * d:], dtype=np.float32)
    assert np.array_equal(bbox[:d], pts.min(axis=0)), "minima"
    assert np.array_equal(bbox[d:], pts.max(axis=0)), "maxima"
    t.close()
