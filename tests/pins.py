"""Independent checks that pin the CPU oracle (oracle/knn_oracle.c).

None of these re-types the oracle's formulas: each compares against exact
integer arithmetic, FP64 mathematics, third-party library routines, or a
closed-form structural property.  Every check returns (ok, message) so the
planted-mistake tests can assert that a wrong oracle fails at least one.
"""
from __future__ import annotations

import numpy as np

U = 2.0 ** -24


def _rows_equal(a, b):
    bad = np.nonzero(~np.all(a == b, axis=1))[0]
    return bad


def lattice_case(seed: int, n: int, nq: int, d: int, side: int = 6):
    """Integer-coordinate points: every FP32 distance is an exact integer."""
    g = np.random.Generator(np.random.Philox(seed))
    pts = g.integers(0, side, (n, d)).astype(np.float32)
    qs = g.integers(0, side, (nq, d)).astype(np.float32)
    if nq > 2 and n > 2:
        qs[0] = pts[1]           # exact duplicate -> distance 0 ties
        pts[2] = pts[1]
    return pts, qs


def exact_lattice_knn(pts, qs, k, max_d2=np.inf, heap=False):
    """Expected result from int64 arithmetic: K smallest (d2, index) among
    points with key below (bits(max_d2), 0xFFFFFFFF)."""
    P = pts.astype(np.int64)
    Q = qs.astype(np.int64)
    d2 = ((Q[:, None, :] - P[None, :, :]) ** 2).sum(-1)  # exact
    md = np.float32(max_d2)
    md_bits = int(md.view(np.uint32))
    nq, n = d2.shape
    out_i = np.full((nq, k), -1, np.int32)
    out_d = np.full((nq, k), md, np.float32)
    for r in range(nq):
        dist = d2[r].astype(np.float32)
        bits = dist.view(np.uint32).astype(np.uint64)
        keys = (bits << np.uint64(32)) | np.arange(n, dtype=np.uint64)
        limit = (np.uint64(md_bits) << np.uint64(32)) | np.uint64(0xFFFFFFFF)
        keys = np.sort(keys[keys < limit])[:k]
        m = len(keys)
        out_i[r, :m] = (keys & np.uint64(0xFFFFFFFF)).astype(np.int64).astype(np.int32)
        out_d[r, :m] = (keys >> np.uint64(32)).astype(np.uint32).view(np.float32)
    if heap:
        out_i, out_d = out_i[:, ::-1].copy(), out_d[:, ::-1].copy()
    return out_i, out_d


def check_exact(idx, dist, exp_i, exp_d):
    bad_i = _rows_equal(idx, exp_i)
    bad_d = _rows_equal(dist.view(np.uint32), exp_d.view(np.uint32))
    ok = len(bad_i) == 0 and len(bad_d) == 0
    return ok, f"rows with wrong index: {bad_i[:5]}, wrong distance bits: {bad_d[:5]}"


def check_fp64(pts, qs, idx, dist, k):
    """Distances within (D+2)u of the FP64 value, and the index set equal to
    the FP64 top-K except inside the rounding band around the K-th distance."""
    d = pts.shape[1]
    tol = (d + 2) * U
    P = pts.astype(np.float64)
    Q = qs.astype(np.float64)
    worst_rel = 0.0
    for r in range(qs.shape[0]):
        d64 = ((Q[r] - P) ** 2).sum(-1)
        got = idx[r][idx[r] >= 0]
        if len(got) == 0:
            continue
        rel = np.abs(dist[r][: len(got)].astype(np.float64) - d64[got]) / np.maximum(d64[got], 1e-300)
        worst_rel = max(worst_rel, float(rel.max()))
        if rel.max() > tol:
            return False, f"query {r}: distance error {rel.max():.3e} > {(d + 2)}u", worst_rel
        kth = np.partition(d64, min(k, len(d64)) - 1)[min(k, len(d64)) - 1]
        if np.any(d64[got] > kth * (1 + 2 * tol) + 1e-300):
            return False, f"query {r}: neighbour beyond the FP64 K-th distance", worst_rel
        must = np.nonzero(d64 < kth * (1 - 2 * tol))[0]
        if not set(must.tolist()) <= set(got.tolist()):
            return False, f"query {r}: missed a neighbour well inside the FP64 K-th distance", worst_rel
    return True, "", worst_rel


def check_f32_sequential(pts, qs, idx, dist):
    """Bits equal numpy float32 step-by-step evaluation (each op rounded once):
    catches FMA contraction or re-association."""
    for r in range(qs.shape[0]):
        sel = idx[r][idx[r] >= 0]
        if len(sel) == 0:
            continue
        q = qs[r].astype(np.float32)
        p = pts[sel].astype(np.float32)
        acc = np.zeros(len(sel), np.float32)
        for j in range(pts.shape[1]):
            t = np.subtract(q[j], p[:, j], dtype=np.float32)
            acc = np.add(acc, np.multiply(t, t, dtype=np.float32), dtype=np.float32)
        if not np.array_equal(acc.view(np.uint32), dist[r][: len(sel)].view(np.uint32)):
            return False, f"query {r}: distance bits differ from sequential FP32"
    return True, ""


def check_sklearn(pts, qs, idx, dist, k):
    """Agreement with scikit-learn's exact brute-force kNN (FP64)."""
    from sklearn.neighbors import NearestNeighbors
    nn = NearestNeighbors(n_neighbors=k, algorithm="brute").fit(pts.astype(np.float64))
    d_sk, i_sk = nn.kneighbors(qs.astype(np.float64))
    d = pts.shape[1]
    tol = 2 * (d + 2) * U
    for r in range(qs.shape[0]):
        kth = d_sk[r, -1] ** 2
        inner = set(i_sk[r][d_sk[r] ** 2 < kth * (1 - tol)].tolist())
        if not inner <= set(idx[r].tolist()):
            return False, f"query {r}: sklearn neighbour missing"
        if np.any(dist[r].astype(np.float64) > kth * (1 + tol) + 1e-30):
            return False, f"query {r}: distance beyond sklearn's K-th"
    return True, ""


def check_scipy_kdtree(pts, qs, idx, k):
    from scipy.spatial import cKDTree
    t = cKDTree(pts.astype(np.float64))
    d_sp, i_sp = t.query(qs.astype(np.float64), k=k)
    i_sp = np.asarray(i_sp).reshape(len(qs), k)
    d_sp = np.asarray(d_sp).reshape(len(qs), k)
    tol = 2 * (pts.shape[1] + 2) * U
    for r in range(len(qs)):
        kth = d_sp[r, -1] ** 2
        inner = set(i_sp[r][d_sp[r] ** 2 < kth * (1 - tol)].tolist())
        if not inner <= set(idx[r].tolist()):
            return False, f"query {r}: scipy cKDTree neighbour missing"
    return True, ""


def check_heap_layout(raw_idx, raw_dist, arr_idx, arr_dist):
    """Raw heap rows are valid max-heaps on the key and hold the array-mode set."""
    k = raw_idx.shape[1]
    for r in range(raw_idx.shape[0]):
        keys = (raw_dist[r].view(np.uint32).astype(np.uint64) << np.uint64(32)) | \
            raw_idx[r].view(np.uint32).astype(np.uint64)
        for i in range(k):
            for c in (2 * i + 1, 2 * i + 2):
                if c < k and keys[c] > keys[i]:
                    return False, f"row {r}: heap property broken at {i}->{c}"
        if sorted(raw_idx[r].tolist()) != sorted(arr_idx[r].tolist()):
            return False, f"row {r}: heap set differs from array set"
    return True, ""


def _ordered(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    neg = (u >> np.uint64(31)) == 1
    return np.where(neg, (~u) & np.uint64(0xFFFFFFFF), u | np.uint64(0x80000000))


def subtree_members(leaf, node_size, max_children, node, node_depth):
    level = int(np.floor(np.log2(node + 1)))
    down = node_depth - level
    leftmost = ((node + 1) << down) - 1
    lo = (leftmost - node_size) * max_children
    hi = lo + (max_children << down)
    m = leaf[lo:hi]
    return m[m >= 0]


def check_kd_structure(pts, info, nodes, leaf, split_dims):
    """Closed-form kd properties: leaves partition [0,N) with -1 suffix
    padding; every split node sends ceil(n/2) points left; its split value is
    the maximum left coordinate (the rank ceil(n/2)-1 order statistic); right
    coordinates are >= it; equal coordinates are split by index; leaf buckets
    are ordered by (last split coordinate, index)."""
    n = pts.shape[0]
    ns, mc, nd = info["node_size"], info["max_children"], info["node_depth"]
    got = leaf[leaf >= 0]
    if len(got) != n or not np.array_equal(np.sort(got), np.arange(n)):
        return False, "leaves are not a permutation of the points"
    for b in range(info["index_size"] // mc):
        bucket = leaf[b * mc:(b + 1) * mc]
        k = int((bucket >= 0).sum())
        if not np.all(bucket[:k] >= 0) or not np.all(bucket[k:] == -1):
            return False, f"bucket {b}: padding is not a suffix"
        if k > 1:
            dim = (nd - 1) % split_dims
            key = (_ordered(pts[bucket[:k], dim]) << np.uint64(32)) | bucket[:k].astype(np.uint64)
            if np.any(key[1:] <= key[:-1]):
                return False, f"bucket {b}: not in key order"
    for node in range(ns):
        level = int(np.floor(np.log2(node + 1)))
        dim = level % split_dims
        L = subtree_members(leaf, ns, mc, 2 * node + 1, nd)
        R = subtree_members(leaf, ns, mc, 2 * node + 2, nd)
        tot = len(L) + len(R)
        if len(L) != (tot + 1) // 2:
            return False, f"node {node}: left holds {len(L)} of {tot}"
        if len(L) == 0:
            continue
        lo = _ordered(pts[L, dim])
        sv = _ordered(np.float32(nodes[node]))
        if lo.max() != sv:
            return False, f"node {node}: split is not the left maximum"
        if len(R):
            ro = _ordered(pts[R, dim])
            if ro.min() < sv:
                return False, f"node {node}: a right point lies below the split"
            tie_l = L[lo == sv]
            tie_r = R[ro == sv]
            if len(tie_r) and tie_l.max() > tie_r.min():
                return False, f"node {node}: ties not split by index"
    return True, ""


def check_ball_structure(pts, info, nodes, dims, leaf):
    """Ball properties: children hold ceil/floor halves split at the median of
    the node's dim; the dim maximises the FP64 sum of squared deviations from
    the node centroid; each child centroid is the FP64 mean within one FP32
    ulp; each child radius bounds the exact (FP64) distances of its points."""
    d = pts.shape[1]
    ns, mc, nd = info["node_size"], info["max_children"], info["node_depth"]
    got = leaf[leaf >= 0]
    if len(got) != pts.shape[0] or not np.array_equal(np.sort(got), np.arange(pts.shape[0])):
        return False, "leaves are not a permutation of the points"
    P64 = pts.astype(np.float64)
    nodes = nodes.reshape(ns, 2 * d + 2)
    for node in range(ns):
        M = subtree_members(leaf, ns, mc, node, nd) if node else got
        L = subtree_members(leaf, ns, mc, 2 * node + 1, nd)
        R = subtree_members(leaf, ns, mc, 2 * node + 2, nd)
        if len(L) != (len(M) + 1) // 2 or len(L) + len(R) != len(M):
            return False, f"node {node}: bad split sizes"
        if node == 0:
            c = P64[M].mean(0).astype(np.float32)
        else:
            par = (node - 1) // 2
            off = 0 if node % 2 == 1 else d + 1
            c = nodes[par, off:off + d]
        if len(M):
            var = ((P64[M] - c.astype(np.float64)) ** 2).sum(0)
            best = int(np.argmax(var)) if var.max() > 0 else 0
            if dims[node] != best and var[dims[node]] < var.max() * (1 - 1e-12):
                return False, f"node {node}: dim {dims[node]} is not the max-deviation dim {best}"
            dm = dims[node]
            if len(L) and len(R) and _ordered(pts[L, dm]).max() > _ordered(pts[R, dm]).min():
                return False, f"node {node}: children overlap on the split dim"
        for child, off in ((L, 0), (R, d + 1)):
            cc = nodes[node, off:off + d]
            r = nodes[node, off + d]
            if len(child) == 0:
                continue
            mean = P64[child].mean(0)
            ulp = np.spacing(np.abs(mean.astype(np.float32))).astype(np.float64)
            if np.any(np.abs(cc.astype(np.float64) - mean) > ulp):
                return False, f"node {node}: child centroid off the FP64 mean by more than 1 ulp"
            rmax = np.sqrt(((P64[child] - cc.astype(np.float64)) ** 2).sum(1).max())
            if r < rmax:
                return False, f"node {node}: radius {r} below the exact radius {rmax}"
    return True, ""


# Tree sizing written out by hand from the reference's constructors
# (src/KdTree.cuh:56-62, src/BallTree.cuh:123-128): with L = log2Ceil(n)
# (src/KnnTrees.cuh:84-93: 1 for n <= 2, else ceil(log2 n)), kd trees clamp
# depthSkips to L - 1 and ball trees to [1, L - 1] (the upper clamp wins when
# L = 1).  Each row is (n, requested skips) -> (kd tuple, ball tuple) with
# tuple = (effective skips, node_depth, max_children, node_size, index_size).
# The values are literals, not evaluations of the formula, so a sizing rule
# shared by the oracle and the library (a dropped clamp, an off-by-one
# log2Ceil) disagrees with at least one row.
SIZING_TABLE = {
    (1, 0): ((0, 1, 1, 1, 2), (0, 1, 1, 1, 2)),
    (1, 1): ((0, 1, 1, 1, 2), (0, 1, 1, 1, 2)),
    (1, 3): ((0, 1, 1, 1, 2), (0, 1, 1, 1, 2)),
    (1, 40): ((0, 1, 1, 1, 2), (0, 1, 1, 1, 2)),
    (2, 0): ((0, 1, 1, 1, 2), (0, 1, 1, 1, 2)),
    (2, 1): ((0, 1, 1, 1, 2), (0, 1, 1, 1, 2)),
    (2, 3): ((0, 1, 1, 1, 2), (0, 1, 1, 1, 2)),
    (2, 40): ((0, 1, 1, 1, 2), (0, 1, 1, 1, 2)),
    (3, 0): ((0, 2, 1, 3, 4), (1, 1, 2, 1, 4)),
    (3, 1): ((1, 1, 2, 1, 4), (1, 1, 2, 1, 4)),
    (3, 3): ((1, 1, 2, 1, 4), (1, 1, 2, 1, 4)),
    (3, 40): ((1, 1, 2, 1, 4), (1, 1, 2, 1, 4)),
    (4, 0): ((0, 2, 1, 3, 4), (1, 1, 2, 1, 4)),
    (4, 1): ((1, 1, 2, 1, 4), (1, 1, 2, 1, 4)),
    (4, 3): ((1, 1, 2, 1, 4), (1, 1, 2, 1, 4)),
    (4, 40): ((1, 1, 2, 1, 4), (1, 1, 2, 1, 4)),
    (5, 0): ((0, 3, 1, 7, 8), (1, 2, 2, 3, 8)),
    (5, 1): ((1, 2, 2, 3, 8), (1, 2, 2, 3, 8)),
    (5, 3): ((2, 1, 4, 1, 8), (2, 1, 4, 1, 8)),
    (5, 40): ((2, 1, 4, 1, 8), (2, 1, 4, 1, 8)),
    (65536, 0): ((0, 16, 1, 65535, 65536), (1, 15, 2, 32767, 65536)),
    (65536, 1): ((1, 15, 2, 32767, 65536), (1, 15, 2, 32767, 65536)),
    (65536, 3): ((3, 13, 8, 8191, 65536), (3, 13, 8, 8191, 65536)),
    (65536, 40): ((15, 1, 32768, 1, 65536), (15, 1, 32768, 1, 65536)),
    (65537, 0): ((0, 17, 1, 131071, 131072), (1, 16, 2, 65535, 131072)),
    (65537, 1): ((1, 16, 2, 65535, 131072), (1, 16, 2, 65535, 131072)),
    (65537, 3): ((3, 14, 8, 16383, 131072), (3, 14, 8, 16383, 131072)),
    (65537, 40): ((16, 1, 65536, 1, 131072), (16, 1, 65536, 1, 131072)),
    # the real-size 3-D workload: kd with depthSkips 0, ball with the default 1
    (16777216, 0): ((0, 24, 1, 16777215, 16777216), (1, 23, 2, 8388607, 16777216)),
    (16777216, 1): ((1, 23, 2, 8388607, 16777216), (1, 23, 2, 8388607, 16777216)),
}
SIZING_FIELDS = ("skips", "node_depth", "max_children", "node_size", "index_size")


def check_sizing(info: dict, n: int, skips: int, ball: bool):
    """Compares a tree's sizing (oracle names: skips; library: depth_skips)
    with the hand-written row for (n, skips)."""
    want = SIZING_TABLE[(n, skips)][1 if ball else 0]
    got = (info.get("skips", info.get("depth_skips")), info["node_depth"], info["max_children"],
           info["node_size"], info["index_size"])
    return got == want, f"n={n} skips={skips} ball={ball}: got {got}, want {want}"
