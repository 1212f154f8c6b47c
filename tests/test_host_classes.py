"""The host classes of include/KnnTrees.hpp (BruteForce<D>, KdTree<D>,
BallTree<D>: src/BruteForce.cuh:8-46, src/KdTree.cuh:10-160,
src/BallTree.cuh:10-240) on the CPU, against the oracle.

tests/cpp/host_classes_test.cpp is compiled with g++ alone (the host classes
need neither HIP nor the library), once at the baseline ISA and once with
-march=native -ffp-contract=fast, which would fuse the distance loop into
FMAs if the header did not forbid it.  The program checks that every tree
row equals brute force's; here brute force's rows are compared with the
oracle's bit for bit, the kd trees' split values and buckets with the
oracle's export, and the ball tree with the structural pins.
"""
import os
import subprocess

import numpy as np
import pytest

from tests import pins

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "host_classes_test.cpp")


@pytest.fixture(scope="module", params=["baseline", "native-fast-contract"])
def binary(request, tmp_path_factory):
    out = str(tmp_path_factory.mktemp("hostcls") / "host_classes_test")
    flags = ["-O2"] if request.param == "baseline" else ["-O3", "-march=native", "-ffp-contract=fast"]
    cmd = ["g++", "-std=c++17", *flags, "-I", os.path.join(ROOT, "include"), SRC, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def _case(d, seed):
    g = np.random.Generator(np.random.Philox(seed))
    n, nq = 3001, 257
    pts = (g.standard_normal((n, d)) * np.exp(g.uniform(-1, 1, d))).astype(np.float32)
    pts[:40, 0] = 0.5                       # ties on the first split dimension
    pts[100] = pts[101]                     # duplicate points: index tie-break
    qs = (g.standard_normal((nq, d)) * 1.3).astype(np.float32)
    qs[0] = pts[101]
    qs[1] = np.round(qs[1])
    return pts, qs


@pytest.mark.parametrize("d", [3, 8])
def test_host_classes_match_oracle(binary, oracle, tmp_path, d):
    pts, qs = _case(d, 40 + d)
    fin = np.float32(0.25 * d)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(np.array([d, len(pts), len(qs)], np.uint32).tobytes())
        f.write(np.array([fin], np.float32).tobytes())
        f.write(pts.tobytes())
        f.write(qs.tobytes())
    r = subprocess.run([binary, str(inp), str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    buf = open(out, "rb").read()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(buf, dtype=dtype, count=count, offset=off)
        off += 4 * count
        return a

    nq = len(qs)
    for md in (np.inf, fin, -1.0):
        for k, heap in ((1, False), (8, False), (8, True), (16, True), (33, False)):
            gi = take(np.int32, nq * k).reshape(nq, k)
            gd = take(np.float32, nq * k).reshape(nq, k)
            wi, wd = oracle.bf_knn(pts, qs, k, max_d2=md, heap=heap)
            assert np.array_equal(gi, wi), f"d={d} k={k} heap={heap} md={md}: indices"
            assert np.array_equal(gd.view(np.uint32), wd.view(np.uint32)), f"d={d} k={k} heap={heap} md={md}"
    for skips in (0, 2):
        o = oracle.KdTree(pts, skips, d)
        info = o.info()
        nodes = take(np.float32, info["node_size"])
        leaf = take(np.int32, info["index_size"])
        onodes, oleaf = o.export()
        assert np.array_equal(nodes.view(np.uint32), onodes.view(np.uint32)), f"kd skips={skips}: split values"
        assert np.array_equal(leaf, oleaf), f"kd skips={skips}: buckets"
    ob = oracle.BallTree(pts, 1)
    info = ob.info()
    ns = info["node_size"]
    nodes = take(np.float32, ns * (2 * d + 2))
    dims = take(np.uint32, ns)
    leaf = take(np.int32, info["index_size"])
    assert off == len(buf)
    ok, msg = pins.check_ball_structure(pts, info, nodes, dims, leaf)
    assert ok, msg
    _, odims, oleaf = ob.export()
    assert np.array_equal(dims, odims)
    assert np.array_equal(np.sort(leaf[leaf >= 0]), np.arange(len(pts)))
