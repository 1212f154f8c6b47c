"""CPU tests: the oracle against independent pins (no GPU needed).

The reference ships no golden vectors or known-answer tests (its only
correctness test, tests/FunctionalityTests.cu:35-107, set-compares unseeded
runs), so the oracle is pinned by exact integer lattices, FP64 mathematics,
scikit-learn / scipy routines, closed-form tree properties and a committed
exact-arithmetic fixture; planted mistakes must fail at least one pin.
"""
import os

import numpy as np
import pytest

from tests import pins

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "lattice_d3.npz")


@pytest.mark.parametrize("d,k", [(1, 3), (2, 5), (3, 8), (4, 10), (7, 17)])
@pytest.mark.parametrize("heap", [False, True])
def test_bf_lattice_exact(oracle, d, k, heap):
    pts, qs = pins.lattice_case(10 + d, 300, 60, d)
    for md in (np.inf, 4.0, 0.0, 1.0, 9.0):
        i, dd = oracle.bf_knn(pts, qs, k, max_d2=md, heap=heap)
        ei, ed = pins.exact_lattice_knn(pts, qs, k, md, heap)
        ok, msg = pins.check_exact(i, dd, ei, ed)
        assert ok, f"max_d2={md}: {msg}"


def test_bf_k_larger_than_n_and_empty(oracle):
    pts, qs = pins.lattice_case(3, 5, 9, 3)
    i, d = oracle.bf_knn(pts, qs, 12)
    ei, ed = pins.exact_lattice_knn(pts, qs, 12)
    assert pins.check_exact(i, d, ei, ed)[0]
    assert (i[:, 5:] == -1).all() and np.isinf(d[:, 5:]).all()
    i, d = oracle.bf_knn(np.zeros((0, 3), np.float32), qs, 4, max_d2=2.5)
    assert (i == -1).all() and (d == np.float32(2.5)).all()


@pytest.mark.parametrize("md", [-1.0, -0.0, float("nan")])
def test_bf_sign_bit_and_nan_radius(oracle, md):
    """A negative or -0.0 maxDistance2 has the sign bit set, so every key is
    below the empty key and brute force accepts all points; empty slots keep
    the given bits (src/BruteForce.cuh:31, 37-42)."""
    pts, qs = pins.lattice_case(4, 40, 10, 3)
    i, d = oracle.bf_knn(pts, qs, 45, max_d2=md)
    ei, ed = pins.exact_lattice_knn(pts, qs, 45, md)
    assert pins.check_exact(i, d, ei, ed)[0]
    assert (i[:, :40] >= 0).all()
    want = np.float32(md).view(np.uint32)
    assert (d[:, 40:].view(np.uint32) == want).all()


def test_golden_fixture(oracle):
    g = np.load(GOLDEN)
    for case in range(int(g["ncases"])):
        md = float(g[f"maxd2_{case}"])
        heap = bool(g[f"heap_{case}"])
        k = int(g[f"k_{case}"])
        i, d = oracle.bf_knn(g["points"], g["queries"], k, max_d2=md, heap=heap)
        ok, msg = pins.check_exact(i, d, g[f"idx_{case}"], g[f"dist_{case}"])
        assert ok, f"case {case}: {msg}"


@pytest.mark.parametrize("d", [3, 8, 16])
def test_bf_fp64_sklearn_scipy(oracle, d):
    g = np.random.Generator(np.random.Philox(d))
    pts = (g.standard_normal((2000, d)) * np.exp(g.uniform(-2, 2, d))).astype(np.float32)
    qs = (g.standard_normal((150, d)) * np.exp(g.uniform(-2, 2, d))).astype(np.float32)
    k = 9
    i, dist = oracle.bf_knn(pts, qs, k)
    ok, msg, worst = pins.check_fp64(pts, qs, i, dist, k)
    assert ok, msg
    assert worst <= (d + 2) * pins.U
    assert pins.check_sklearn(pts, qs, i, dist, k)[0]
    assert pins.check_scipy_kdtree(pts, qs, i, k)[0]
    assert pins.check_f32_sequential(pts, qs, i, dist)[0]


def test_heap_layout(oracle):
    pts, qs = pins.lattice_case(5, 400, 50, 3, side=9)
    oracle.set_raw_heap(True)
    try:
        ri, rd = oracle.bf_knn(pts, qs, 11, heap=True)
    finally:
        oracle.set_raw_heap(False)
    ai, ad = oracle.bf_knn(pts, qs, 11, heap=False)
    ok, msg = pins.check_heap_layout(ri, rd, ai, ad)
    assert ok, msg
    ci, cd = oracle.bf_knn(pts, qs, 11, heap=True)
    assert (ci == ai[:, ::-1]).all() and (cd == ad[:, ::-1]).all()


@pytest.mark.parametrize("skips,sd", [(0, 3), (1, 3), (3, 3), (2, 2), (0, 1)])
def test_kd_structure_and_query(oracle, skips, sd):
    g = np.random.Generator(np.random.Philox(skips * 7 + sd))
    pts = g.random((777, 3), dtype=np.float32)
    pts[:40, 0] = 0.5  # many equal coordinates on dim 0
    pts[50] = pts[51]
    pts[60, 1] = -0.0
    pts[61, 1] = 0.0
    t = oracle.KdTree(pts, skips, sd)
    nodes, leaf = t.export()
    ok, msg = pins.check_kd_structure(pts, t.info(), nodes, leaf, sd)
    assert ok, msg
    qs = g.random((120, 3), dtype=np.float32)
    qs[0] = pts[51]
    for md in (np.inf, 0.01, 0.0, -1.0):
        for heap in (False, True):
            bi, bd = oracle.bf_knn(pts, qs, 10, max_d2=md, heap=heap)
            ti, td = t.knn(qs, 10, max_d2=md, heap=heap)
            assert (bi == ti).all() and (bd.view(np.uint32) == td.view(np.uint32)).all()


@pytest.mark.parametrize("skips", [0, 1, 2, 4])
def test_ball_structure_and_query(oracle, skips):
    g = np.random.Generator(np.random.Philox(100 + skips))
    pts = (g.standard_normal((600, 4)) * np.array([3, 1, 0.2, 2])).astype(np.float32)
    pts[10] = pts[11]
    b = oracle.BallTree(pts, skips)
    nodes, dims, leaf = b.export()
    ok, msg = pins.check_ball_structure(pts, b.info(), nodes, dims, leaf)
    assert ok, msg
    qs = (g.standard_normal((100, 4)) * 2).astype(np.float32)
    qs[0] = pts[11]
    for md in (np.inf, 0.5, 0.0, -0.0):
        for heap in (False, True):
            bi, bd = oracle.bf_knn(pts, qs, 7, max_d2=md, heap=heap)
            ti, td = b.knn(qs, 7, max_d2=md, heap=heap)
            assert (bi == ti).all() and (bd.view(np.uint32) == td.view(np.uint32)).all()


def test_ball_bound_is_certified(oracle):
    """The traversal bound never exceeds any computed distance inside the ball."""
    g = np.random.Generator(np.random.Philox(77))
    for d in (2, 3, 8):
        for _ in range(200):
            p = (g.standard_normal((20, d)) * g.uniform(0.01, 100)).astype(np.float32)
            c = p.astype(np.float64).mean(0).astype(np.float32)
            r = max(np.sqrt(max(float(oracle.dist2(c, x)) for x in p)), 0.0)
            r = np.float32(r) * np.float32(1.0 + (d + 8) * 2.0 ** -23) + np.float32(2.0 ** -100)
            q = (c + g.standard_normal(d) * g.uniform(0, 300)).astype(np.float32)
            lb = oracle.ball_bound(q, c, r, d)
            mind = min(float(oracle.dist2(q, x)) for x in p)
            assert lb <= mind


def test_strict_reference_pruning_drops_ties(oracle):
    """The reference trees prune with `bound >= maxDistance2`
    (src/KdTree.cuh:114,130): with maxDistance2 equal to an exact lattice
    distance they can lose a point brute force keeps.  The certified
    non-strict rule used here keeps it."""
    pts, qs = pins.lattice_case(9, 200, 80, 2, side=5)
    diffs = 0
    for md in (1.0, 2.0, 4.0):
        bi, _ = oracle.bf_knn(pts, qs, 6, max_d2=md)
        t = oracle.KdTree(pts, 0)
        si, _ = t.knn(qs, 6, max_d2=md, strict_ref=True)
        ci, _ = t.knn(qs, 6, max_d2=md)
        assert (ci == bi).all()
        diffs += int((si != bi).any(axis=1).sum())
    assert diffs > 0


def test_uniform_c1_sample_against_fp64(oracle):
    from cudaexactknn_amd import datasets
    pts, qs = datasets.uniform_cube(65536, 16384, 3, 0)
    qs = qs[:64]
    i, d = oracle.bf_knn(pts, qs, 8, heap=True)
    i2, d2 = oracle.bf_knn(pts, qs, 8, heap=False)
    assert (i == i2[:, ::-1]).all()
    ok, msg, _ = pins.check_fp64(pts, qs, i2, d2, 8)
    assert ok, msg
    kd = oracle.KdTree(pts, 0)
    ki, kd_ = kd.knn(qs, 8, heap=True)
    assert (ki == i).all()


# --- planted mistakes: each must be caught by at least one pin ---------------

def _pins_all(oracle):
    fails = []
    pts, qs = pins.lattice_case(21, 250, 40, 3)
    for heap in (False, True):
        i, d = oracle.bf_knn(pts, qs, 9, heap=heap)
        ei, ed = pins.exact_lattice_knn(pts, qs, 9, np.inf, heap)
        if not pins.check_exact(i, d, ei, ed)[0]:
            fails.append("lattice")
    oracle.set_raw_heap(True)
    ri, rd = oracle.bf_knn(pts, qs, 9, heap=True)
    oracle.set_raw_heap(False)
    ai, ad = oracle.bf_knn(pts, qs, 9)
    if not pins.check_heap_layout(ri, rd, ai, ad)[0]:
        fails.append("heap")
    g = np.random.Generator(np.random.Philox(5))
    rp = g.random((1500, 3), dtype=np.float32)
    rq = g.random((200, 3), dtype=np.float32)
    bi, bd = oracle.bf_knn(rp, rq, 8)
    if not pins.check_fp64(rp, rq, bi, bd, 8)[0]:
        fails.append("fp64")
    t = oracle.KdTree(rp, 0)
    nodes, leaf = t.export()
    if not pins.check_kd_structure(rp, t.info(), nodes, leaf, 3)[0]:
        fails.append("kd-structure")
    oracle.set_mutation(0)
    ref_i, _ = oracle.bf_knn(rp, rq, 8)
    return fails, rp, rq, ref_i


@pytest.mark.parametrize("mutation", [1, 2, 3, 4, 5, 6, 7])
def test_planted_mistakes_are_caught(oracle, mutation):
    oracle.set_mutation(mutation)
    try:
        fails, rp, rq, ref_i = _pins_all(oracle)
        oracle.set_mutation(mutation)
        t = oracle.KdTree(rp, 0)
        ti, _ = t.knn(rq, 8)
        b = oracle.BallTree(rp, 1)
        bi, _ = b.knn(rq, 8)
        nodes, dims, leaf = b.export()
        oracle.set_mutation(0)
        if (ti != ref_i).any():
            fails.append("kd==bf")
        if (bi != ref_i).any():
            fails.append("ball==bf")
        if not pins.check_ball_structure(rp, b.info(), nodes, dims, leaf)[0]:
            fails.append("ball-structure")
    finally:
        oracle.set_mutation(0)
    assert fails, f"mutation {mutation} passed every pin"


def test_faithful_oracle_passes_all_pins(oracle):
    fails, *_ = _pins_all(oracle)
    assert fails == []


def test_mlp_oracle_gradient_matches_finite_differences():
    """The FP64 MLP oracle's analytic gradient equals central differences."""
    from oracle import mlp
    g = np.random.Generator(np.random.Philox(3))
    p = g.standard_normal(mlp.N_PARAMS) * 0.2
    X = g.standard_normal((40, mlp.F_IN))
    T = g.standard_normal(40)
    grad, _ = mlp.gradient(p, X, T)
    idx = g.choice(mlp.N_PARAMS, 200, replace=False)
    idx = np.concatenate([idx, [mlp.N_PARAMS - 1, 0, 12 * 64, 12 * 64 + 64]])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (mlp.loss(p + e, X, T) - mlp.loss(p - e, X, T)) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, grad[i])
