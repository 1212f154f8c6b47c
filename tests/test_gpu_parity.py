"""GPU parity tests: the HIP library through its C ABI against the CPU oracle.

Bar: indices and distance bits identical to the oracle's BruteForce (which
restates src/BruteForce.cuh:27-44) for every structure, option and edge case;
tree structures identical to the oracle's (kd: split values and leaf buckets
bit for bit; ball: dims and leaves, centroids within one FP32 ulp).
"""
import numpy as np
import pytest

import cudaexactknn_amd as knn
from cudaexactknn_amd import datasets
from tests import pins

pytestmark = pytest.mark.gpu


def dev(x, cuda):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to(cuda)


def run(struct, qs, k, cuda, md=np.inf, **opts):
    import torch
    q = dev(qs, cuda)
    ri = torch.empty((len(qs), k), dtype=torch.int32, device=cuda)
    rd = torch.empty((len(qs), k), dtype=torch.float32, device=cuda)
    struct.batchKnn(knn.Query(q, len(qs), md, ri, rd), k, **opts)
    torch.cuda.synchronize()
    return ri.cpu().numpy(), rd.cpu().numpy()


def same(a, b, what=""):
    (ai, ad), (bi, bd) = a, b
    bad = np.nonzero(~(np.all(ai == bi, axis=1) & np.all(ad.view(np.uint32) == bd.view(np.uint32), axis=1)))[0]
    assert len(bad) == 0, f"{what}: {len(bad)} rows differ, first {bad[:5]}: got {ai[bad[0]]} want {bi[bad[0]]}"


BF_VARIANTS = [dict(USE_HEAP_TRACKING=h, USE_TRACKING_SHMEM=s) for h in (False, True) for s in (False, True)]
TREE_VARIANTS = [dict(USE_HEAP_TRACKING=h, USE_TRACKING_SHMEM=s, USE_STACK_SHMEM=t)
                 for h in (False, True) for s in (False, True) for t in (False, True)]


@pytest.mark.parametrize("d", [2, 3, 4, 5, 8, 16, 33, 128])
def test_bf_dims_and_variants(oracle, cuda, d):
    g = np.random.Generator(np.random.Philox(d))
    pts = (g.standard_normal((1500, d)) * np.exp(g.uniform(-1, 1, d))).astype(np.float32)
    qs = (g.standard_normal((300, d)) * 1.5).astype(np.float32)
    bf = knn.BruteForceGpu(dev(pts, cuda))
    for k in (1, 8, 10, 32, 40):
        for v in BF_VARIANTS:
            want = oracle.bf_knn(pts, qs, k, heap=v["USE_HEAP_TRACKING"])
            same(run(bf, qs, k, cuda, BLOCK_SIZE=64, **v), want, f"d={d} k={k} {v}")


@pytest.mark.parametrize("d", [3, 8, 32, 64, 128, 7])
def test_bf_split_scan_matches_oracle(oracle, cuda, d):
    """Few queries and many points: the scan is split across workgroups
    (partial top-K per point chunk, then a K-way merge)."""
    g = np.random.Generator(np.random.Philox(70 + d))
    pts = (g.standard_normal((30000, d)) * np.exp(g.uniform(-1, 1, d))).astype(np.float32)
    pts[100] = pts[20000]
    qs = (g.standard_normal((200, d)) * 1.3).astype(np.float32)
    qs[0] = pts[20000]
    bf = knn.BruteForceGpu(dev(pts, cuda))
    for k in (1, 16, 33):
        for v in BF_VARIANTS:
            for md in (np.inf, float(np.float32(0.5 * d))):
                want = oracle.bf_knn(pts, qs, k, max_d2=md, heap=v["USE_HEAP_TRACKING"])
                same(run(bf, qs, k, cuda, md=md, BLOCK_SIZE=128, **v), want, f"split d={d} k={k} md={md} {v}")


def test_bf_lattice_and_radius_edges(oracle, cuda):
    pts, qs = pins.lattice_case(33, 500, 90, 3)
    bf = knn.BruteForceGpu(dev(pts, cuda))
    for md in (np.inf, 0.0, -0.0, -2.0, 1.0, 4.0, float("nan"), 1e-40):
        for v in BF_VARIANTS:
            for k in (3, 16):
                ei, ed = pins.exact_lattice_knn(pts, qs, k, md, v["USE_HEAP_TRACKING"])
                same(run(bf, qs, k, cuda, md=md, BLOCK_SIZE=64, **v), (ei, ed), f"md={md} {v}")


def test_golden_fixture_all_structures(cuda):
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "lattice_d3.npz"))
    p = dev(g["points"], cuda)
    structs = [knn.BruteForceGpu(p), knn.KdTreeGpu(p, depth_skips=0), knn.KdTreeGpu(p, depth_skips=2),
               knn.BallTreeGpu(p, depth_skips=1), knn.BallTreeGpu(p, depth_skips=3)]
    for case in range(int(g["ncases"])):
        md, heap, k = float(g[f"maxd2_{case}"]), bool(g[f"heap_{case}"]), int(g[f"k_{case}"])
        for s in structs:
            got = run(s, g["queries"], k, cuda, md=md, BLOCK_SIZE=32 if k > 32 else 0, USE_HEAP_TRACKING=heap)
            same(got, (g[f"idx_{case}"], g[f"dist_{case}"]), f"case {case} {type(s).__name__}")


def test_bf_tiny_sizes(oracle, cuda):
    for n in (0, 1, 2, 3, 7):
        pts = np.arange(n * 3, dtype=np.float32).reshape(n, 3) * 0.25
        qs = np.random.Generator(np.random.Philox(n)).random((5, 3), dtype=np.float32)
        if n == 0:
            import torch
            bf = knn.BruteForceGpu(torch.empty((0, 3), dtype=torch.float32, device=cuda))
        else:
            bf = knn.BruteForceGpu(dev(pts, cuda))
        for k in (1, 4, 9):
            same(run(bf, qs, k, cuda, md=3.0), oracle.bf_knn(pts.reshape(n, 3), qs, k, max_d2=3.0), f"n={n} k={k}")


def test_bf_special_values(oracle, cuda):
    g = np.random.Generator(np.random.Philox(99))
    pts = g.standard_normal((400, 4)).astype(np.float32)
    pts[0] = [-0.0, 0.0, -0.0, 0.0]
    pts[1] = [1e-41, -1e-42, 3e-39, 0.0]      # denormals
    pts[2] = [3e38, -3e38, 1e30, 0.0]         # distances overflow to +inf
    pts[3] = pts[4]
    qs = g.standard_normal((64, 4)).astype(np.float32)
    qs[0] = pts[1]
    qs[1] = [-3e38, 3e38, 0.0, 0.0]
    bf = knn.BruteForceGpu(dev(pts, cuda))
    for v in BF_VARIANTS:
        same(run(bf, qs, 12, cuda, **v), oracle.bf_knn(pts, qs, 12, heap=v["USE_HEAP_TRACKING"]), str(v))
        same(run(bf, qs, 450, cuda, BLOCK_SIZE=16, **v), oracle.bf_knn(pts, qs, 450, heap=v["USE_HEAP_TRACKING"]),
             "k>n " + str(v))


@pytest.mark.parametrize("d,sd,skips", [(3, 3, 0), (3, 3, 3), (4, 4, 0), (4, 3, 3), (8, 8, 2), (2, 1, 1),
                                        (5, 5, 1), (6, 2, 0)])
def test_kd_structure_matches_oracle(oracle, cuda, d, sd, skips):
    g = np.random.Generator(np.random.Philox(d * 100 + skips))
    pts = g.random((3001, d), dtype=np.float32)
    pts[:100, 0] = 0.25
    pts[200] = pts[201]
    t = knn.KdTreeGpu(dev(pts, cuda), depth_skips=skips, split_dimensions=sd)
    o = oracle.KdTree(pts, skips, sd)
    info = t.info()
    for key in ("node_depth", "max_children", "node_size", "index_size"):
        assert info[key] == o.info()[key]
    nodes, leaf_idx, leaf_pts, _ = t.export()
    onodes, oleaf = o.export()
    assert np.array_equal(nodes.view(np.uint32), onodes.view(np.uint32))
    assert np.array_equal(leaf_idx, oleaf)
    m = leaf_idx >= 0
    assert np.array_equal(leaf_pts[m], pts[leaf_idx[m]])


@pytest.mark.parametrize("n,d,sd,skips", [(4096, 3, 3, 0), (4097, 3, 3, 0), (50000, 3, 3, 0), (50000, 4, 3, 2),
                                          (50000, 8, 8, 1), (40000, 16, 16, 0), (20000, 30, 30, 0),
                                          (50000, 2, 1, 1), (50000, 2, 2, 0), (70000, 6, 5, 0), (70000, 6, 4, 1)])
def test_kd_structure_global_then_local_levels(oracle, cuda, n, d, sd, skips):
    """Sizes where the build runs some levels over the whole point set and
    finishes each segment inside one workgroup (segments of at most 4096,
    2048 or 1024 points by the number of split lists; 30 lists: every level
    global).  A global level partitions a list only when its dimension
    splits next or when it has waited two levels, so one, two and four or
    more split lists take different partition schedules.  Ties on the first
    coordinate and duplicate points included."""
    g = np.random.Generator(np.random.Philox(n + d * 7 + skips))
    pts = g.random((n, d), dtype=np.float32)
    pts[: n // 8, 0] = 0.5
    pts[n // 2 : n // 2 + 64] = pts[n // 2 + 64]
    t = knn.KdTreeGpu(dev(pts, cuda), depth_skips=skips, split_dimensions=sd)
    o = oracle.KdTree(pts, skips, sd)
    nodes, leaf_idx, leaf_pts, _ = t.export()
    onodes, oleaf = o.export()
    assert np.array_equal(nodes.view(np.uint32), onodes.view(np.uint32))
    assert np.array_equal(leaf_idx, oleaf)
    m = leaf_idx >= 0
    assert np.array_equal(leaf_pts[m], pts[leaf_idx[m]])


@pytest.mark.parametrize("d,skips", [(3, 1), (3, 0), (4, 2), (8, 1), (5, 3)])
def test_ball_structure_matches_oracle(oracle, cuda, d, skips):
    g = np.random.Generator(np.random.Philox(500 + d * 10 + skips))
    pts = (g.standard_normal((2500, d)) * np.exp(g.uniform(-1, 1, d))).astype(np.float32)
    t = knn.BallTreeGpu(dev(pts, cuda), depth_skips=skips)
    o = oracle.BallTree(pts, skips)
    nodes, leaf_idx, _, dims = t.export()
    onodes, odims, oleaf = o.export()
    assert np.array_equal(dims, odims)
    assert np.array_equal(leaf_idx, oleaf)
    a, b = nodes.reshape(onodes.shape), onodes
    ulp = np.spacing(np.abs(b)).astype(np.float64)
    assert np.all(np.abs(a.astype(np.float64) - b) <= 2 * ulp)
    ok, msg = pins.check_ball_structure(pts, t.info(), nodes, dims, leaf_idx)
    assert ok, msg


@pytest.mark.parametrize("n,d,skips", [(4096, 3, 1), (4097, 3, 1), (50000, 3, 1), (50000, 4, 2), (30000, 8, 1),
                                        (20000, 9, 1)])
def test_ball_structure_global_then_local_levels(oracle, cuda, n, d, skips):
    """Sizes where the ball build runs some levels over the whole point set
    and finishes each segment inside one workgroup (segments of at most 2048
    points for D = 2, 3 and 1024 for D = 4 .. 8; D = 9: every level global)."""
    g = np.random.Generator(np.random.Philox(700 + n + d * 10 + skips))
    pts = (g.standard_normal((n, d)) * np.exp(g.uniform(-1, 1, d))).astype(np.float32)
    pts[n // 2 : n // 2 + 40] = pts[n // 2 + 40]
    t = knn.BallTreeGpu(dev(pts, cuda), depth_skips=skips)
    o = oracle.BallTree(pts, skips)
    nodes, leaf_idx, _, dims = t.export()
    onodes, odims, oleaf = o.export()
    assert np.array_equal(dims, odims)
    assert np.array_equal(leaf_idx, oleaf)
    a, b = nodes.reshape(onodes.shape), onodes
    ulp = np.spacing(np.abs(b)).astype(np.float64)
    assert np.all(np.abs(a.astype(np.float64) - b) <= 2 * ulp)
    ok, msg = pins.check_ball_structure(pts, t.info(), nodes, dims, leaf_idx)
    assert ok, msg


@pytest.mark.parametrize("kind", ["kd", "ball"])
@pytest.mark.parametrize("d", [3, 4, 8, 6, 12, 17])
def test_tree_queries_all_variants(oracle, cuda, kind, d):
    """Every option set on both trees.  D = 6, 12 and 17 run the
    wave-cooperative kernels compiled for 8, 16 and 32 dims with the padded
    coordinates zero."""
    g = np.random.Generator(np.random.Philox(1000 + d))
    pts = (g.standard_normal((4000, d)) * np.exp(g.uniform(-1, 1, d))).astype(np.float32)
    pts[7] = pts[8]
    qs = (g.standard_normal((700, d)) * 1.2).astype(np.float32)
    qs[0] = pts[8]
    p = dev(pts, cuda)
    trees = ([knn.KdTreeGpu(p, depth_skips=s) for s in (0, 3)] if kind == "kd"
             else [knn.BallTreeGpu(p, depth_skips=s) for s in (1, 4)])
    for k in (1, 10, 16, 33):
        for md in (np.inf, 0.3):
            want = {h: oracle.bf_knn(pts, qs, k, max_d2=md, heap=h) for h in (False, True)}
            for t in trees:
                for v in TREE_VARIANTS:
                    for span in (0, 0xFFFFFFFF, 2):
                        for per_lane in (False, True):
                            for bs in ((64, 192) if not per_lane else (64,)):
                                got = run(t, qs, k, cuda, md=md, BLOCK_SIZE=bs, leaf_span=span, per_lane=per_lane, **v)
                                same(got, want[v["USE_HEAP_TRACKING"]],
                                     f"{kind} d={d} k={k} md={md} span={span} lane={per_lane} bs={bs} {v}")


@pytest.mark.parametrize("kind", ["kd", "ball"])
@pytest.mark.parametrize("d,sd", [(1, 1), (5, 5), (5, 2), (9, 9), (16, 7), (24, 24), (32, 32)])
def test_padded_dims_traversal(oracle, cuda, kind, d, sd):
    """Dims without a kernel of their own (1, 5, 9, 16, 24, 32) run the
    wave-cooperative kernel of the next padded width (block sizes 64 and
    256): rows equal the oracle and the per-lane traversal (block sizes 16
    and 32, the reference speed sweeps' shapes), for split dims below D too."""
    if kind == "ball" and d == 1:
        pytest.skip("one-dimensional ball trees are refused: their first GPU run took a memory fault whose "
                    "cause is not found yet (test_abi checks the refusal on the CPU)")
    g = np.random.Generator(np.random.Philox(4200 + 10 * d + sd))
    pts = g.random((6000, d), dtype=np.float32)
    pts[11] = pts[12]
    qs = g.random((900, d), dtype=np.float32)
    qs[3] = pts[12]
    p = dev(pts, cuda)
    trees = ([knn.KdTreeGpu(p, split_dimensions=sd, depth_skips=s) for s in (0, 3)] if kind == "kd"
             else [knn.BallTreeGpu(p, depth_skips=s) for s in (1, 2)])
    for k in (5, 16, 32):
        for md in (np.inf, float(np.float32(0.01 * d))):
            want = {h: oracle.bf_knn(pts, qs, k, max_d2=md, heap=h) for h in (False, True)}
            for t in trees:
                for v in TREE_VARIANTS[::3]:
                    for bs in (16, 32, 64, 256):
                        same(run(t, qs, k, cuda, md=md, BLOCK_SIZE=bs, **v), want[v["USE_HEAP_TRACKING"]],
                             f"{kind} d={d} sd={sd} k={k} md={md} bs={bs} {v}")
                same(run(t, qs, k, cuda, md=md, per_lane=True), want[False], f"{kind} d={d} per-lane")


@pytest.mark.parametrize("kind", ["kd", "ball"])
@pytest.mark.parametrize("d", [3, 8])
def test_visit_counting_kernels_are_exact(oracle, cuda, kind, d):
    """The visit-counting traversals are compiled apart from the timed ones;
    they return the same rows, for every buffer kind, and count visits."""
    g = np.random.Generator(np.random.Philox(31 + d))
    pts = g.random((20000, d), dtype=np.float32)
    qs = g.random((3000, d), dtype=np.float32)
    t = knn.KdTreeGpu(dev(pts, cuda)) if kind == "kd" else knn.BallTreeGpu(dev(pts, cuda))
    for k in (8, 16, 32):
        want = {h: oracle.bf_knn(pts, qs, k, heap=h) for h in (False, True)}
        for v in TREE_VARIANTS:
            for per_lane in (False, True):
                got = run(t, qs, k, cuda, count_visits=True, per_lane=per_lane, **v)
                same(got, want[v["USE_HEAP_TRACKING"]], f"{kind} d={d} k={k} counting lane={per_lane} {v}")
                st = t.last_query_stats()
                assert st["node_bytes"] > 0 and st["leaf_slots"] >= len(qs), st


@pytest.mark.parametrize("kind", ["kd", "ball"])
def test_tree_radius_edges_and_ties(oracle, cuda, kind):
    pts, qs = pins.lattice_case(77, 900, 120, 3, side=7)
    p = dev(pts, cuda)
    t = knn.KdTreeGpu(p) if kind == "kd" else knn.BallTreeGpu(p)
    for md in (np.inf, 0.0, -0.0, -5.0, 1.0, 2.0, 3.0, float("nan")):
        for heap in (False, True):
            ei, ed = pins.exact_lattice_knn(pts, qs, 14, md, heap)
            for per_lane in (False, True):
                same(run(t, qs, 14, cuda, md=md, USE_HEAP_TRACKING=heap, per_lane=per_lane), (ei, ed),
                     f"{kind} md={md} heap={heap} lane={per_lane}")


@pytest.mark.parametrize("n", [0, 1, 2, 3, 5, 17])
def test_tree_tiny_sizes(oracle, cuda, n):
    import torch
    g = np.random.Generator(np.random.Philox(n + 3))
    pts = g.random((n, 3), dtype=np.float32)
    qs = g.random((9, 3), dtype=np.float32)
    p = dev(pts, cuda) if n else torch.empty((0, 3), dtype=torch.float32, device=cuda)
    for t in (knn.KdTreeGpu(p), knn.KdTreeGpu(p, depth_skips=2), knn.BallTreeGpu(p), knn.BallTreeGpu(p, depth_skips=0)):
        for k in (1, 4):
            same(run(t, qs, k, cuda, md=2.0), oracle.bf_knn(pts, qs, k, max_d2=2.0), f"n={n} k={k}")


def test_zero_queries_is_a_noop(cuda):
    import torch
    pts = np.random.Generator(np.random.Philox(1)).random((50, 3), dtype=np.float32)
    t = knn.KdTreeGpu(dev(pts, cuda))
    e = torch.empty((0, 3), dtype=torch.float32, device=cuda)
    ri = torch.empty((0, 4), dtype=torch.int32, device=cuda)
    rd = torch.empty((0, 4), dtype=torch.float32, device=cuda)
    t.batchKnn(knn.Query(e, 0, 1.0, ri, rd), 4)


def test_reference_functionality_matrix(oracle, cuda):
    """tests/FunctionalityTests.cu of the reference: D=4, 100,000 points,
    50,000 queries, K=10, maxDistance2=5.9, every CPU and GPU variant; here
    seeded, and compared bit-exact (the reference set-compares)."""
    g = np.random.Generator(np.random.Philox(4))
    pts = g.random((100000, 4), dtype=np.float32)
    qs = g.random((50000, 4), dtype=np.float32)
    md = 5.9
    want = oracle.bf_knn(pts, qs, 10, max_d2=md)
    want_h = oracle.bf_knn(pts, qs, 10, max_d2=md, heap=True)
    p = dev(pts, cuda)
    bf = knn.BruteForceGpu(p)
    same(run(bf, qs, 10, cuda, md=md, BLOCK_SIZE=16), want, "bf array")
    same(run(bf, qs, 10, cuda, md=md, BLOCK_SIZE=16, USE_HEAP_TRACKING=True), want_h, "bf heap")
    same(run(bf, qs, 10, cuda, md=md, BLOCK_SIZE=16, USE_TRACKING_SHMEM=True), want, "bf shmem")
    kd = knn.KdTreeGpu(p, depth_skips=0, split_dimensions=4, block_size=16)
    same(run(kd, qs, 10, cuda, md=md, BLOCK_SIZE=16, USE_HEAP_TRACKING=True, USE_STACK_SHMEM=True), want_h, "kd 4/4")
    kd2 = knn.KdTreeGpu(p, depth_skips=3, split_dimensions=3, block_size=16)
    same(run(kd2, qs, 10, cuda, md=md, BLOCK_SIZE=16, USE_TRACKING_SHMEM=True), want, "kd 4/3")
    ball = knn.BallTreeGpu(p, depth_skips=0, block_size=16)
    same(run(ball, qs, 10, cuda, md=md, BLOCK_SIZE=16, USE_HEAP_TRACKING=True, USE_TRACKING_SHMEM=True,
             USE_STACK_SHMEM=True), want_h, "ball sh/sh")
    ball2 = knn.BallTreeGpu(p, depth_skips=2, block_size=16)
    same(run(ball2, qs, 10, cuda, md=md, BLOCK_SIZE=16), want, "ball reg/reg")
    # CPU oracle trees (the reference's KdTree/BallTree rows)
    same(oracle.KdTree(pts, 3).knn(qs[:5000], 10, max_d2=md), tuple(a[:5000] for a in want), "cpu kd")
    same(oracle.BallTree(pts, 3).knn(qs[:5000], 10, max_d2=md), tuple(a[:5000] for a in want), "cpu ball")


def test_uniform_small_config_bit_exact(oracle, cuda):
    """65,536 uniform points, 16,384 queries, D=3, K=8, heap tracking,
    maxRadius=+inf, seed 0: BruteForce and KdTree (depthSkips=0)."""
    pts, qs = datasets.uniform_cube(65536, 16384, 3, 0)
    want = oracle.bf_knn(pts, qs, 8, heap=True)
    p = dev(pts, cuda)
    same(run(knn.BruteForceGpu(p), qs, 8, cuda, USE_HEAP_TRACKING=True), want, "bf")
    kd = knn.KdTreeGpu(p, depth_skips=0)
    for v in TREE_VARIANTS:
        if v["USE_HEAP_TRACKING"]:
            same(run(kd, qs, 8, cuda, **v), want, f"kd {v}")


@pytest.mark.parametrize("d", [3, 16, 32, 64, 100, 128])
def test_bf16_form_equals_plain(oracle, cuda, d):
    """BF16-MFMA filter + exact rerank returns the plain FP32 result bit for
    bit, on per-dimension-scaled Gaussian data, with ties and near-ties."""
    g = np.random.Generator(np.random.Philox(900 + d))
    scale = np.exp(g.uniform(np.log(0.1), np.log(10.0), d))
    pts = (g.standard_normal((20000, d)) * scale).astype(np.float32)
    qs = (g.standard_normal((700, d)) * scale).astype(np.float32)
    pts[5] = pts[6]                        # duplicate point: exact tie
    qs[0] = pts[6]                         # query on a point: distance 0
    qs[1] = pts[7] + np.float32(1e-6)      # near-coincident
    qs[2] = pts[8] * np.float32(1 + 2.0 ** -20)
    bf = knn.BruteForceGpu(dev(pts, cuda))
    for k in (1, 5, 8, 31, 32):
        for heap in (False, True):
            for md in (np.inf, float(np.float32(np.sum(scale ** 2)))):
                want = oracle.bf_knn(pts, qs, k, max_d2=md, heap=heap)
                got = run(bf, qs, k, cuda, md=md, USE_HEAP_TRACKING=heap, form=knn.FORM_BF16)
                same(got, want, f"bf16 d={d} k={k} heap={heap} md={md}")
    st = bf.last_query_stats()
    assert 0.0 < st["bf16_survivor_fraction"] < 0.5


@pytest.mark.parametrize("d", [16, 32, 128])
def test_bf16_staging_depths_agree(oracle, cuda, d):
    """Every point-tile staging depth (tiles per barrier, up to more tiles
    than a workgroup has threads for their norms) gives the plain result."""
    g = np.random.Generator(np.random.Philox(1300 + d))
    pts = g.standard_normal((20000, d)).astype(np.float32)
    qs = g.standard_normal((300, d)).astype(np.float32)
    qs[0] = pts[11]
    bf = knn.BruteForceGpu(dev(pts, cuda))
    want = oracle.bf_knn(pts, qs, 16)
    for tile in (0, 1, 2, 4, 8, 16, 32):
        same(run(bf, qs, 16, cuda, form=knn.FORM_BF16, tile=tile), want, f"bf16 d={d} tile={tile}")


def test_bf16_form_overflow_falls_back(oracle, cuda):
    """Every point equidistant from the query: all are candidates, the
    candidate buffer overflows and the plain kernel recomputes the query."""
    g = np.random.Generator(np.random.Philox(5))
    d = 32
    v = g.standard_normal((6000, d))
    pts = (v / np.linalg.norm(v, axis=1, keepdims=True) * 4.0).astype(np.float32)
    qs = np.zeros((3, d), np.float32)
    qs[1] = pts[0]
    bf = knn.BruteForceGpu(dev(pts, cuda))
    for k in (4, 32):
        want = oracle.bf_knn(pts, qs, k)
        same(run(bf, qs, k, cuda, form=knn.FORM_BF16), want, f"overflow k={k}")
    assert bf.last_query_stats()["rerun_queries"] >= 1


def test_bf16_form_radius_and_small_sets(oracle, cuda):
    g = np.random.Generator(np.random.Philox(17))
    for n in (1, 31, 33, 100):
        pts = g.standard_normal((n, 40)).astype(np.float32)
        qs = g.standard_normal((50, 40)).astype(np.float32)
        bf = knn.BruteForceGpu(dev(pts, cuda))
        for md in (np.inf, 60.0, 0.0, -1.0):
            for k in (1, 5, 40):
                want = oracle.bf_knn(pts, qs, k, max_d2=md)
                same(run(bf, qs, k, cuda, md=md, form=knn.FORM_BF16), want, f"n={n} md={md} k={k}")


@pytest.mark.parametrize("d", [3, 5])
@pytest.mark.parametrize("k", [1, 8, 32])
@pytest.mark.parametrize("md", [np.inf, 0.004, 0.0, -1.0])
def test_bf_larger_sets_radius_and_tie_edges(oracle, cuda, d, k, md):
    """40,000 points (split scans with partial merges): duplicated points and
    lattice ties around the K-th neighbour, small, zero and negative radii,
    every K-buffer variant."""
    g = np.random.Generator(np.random.Philox(1000 + 10 * d + k))
    n = 40000
    pts = g.random((n, d), dtype=np.float32)
    pts[n // 2:n // 2 + 64] = pts[:64]          # duplicates: index tie-break
    qs = g.random((1500, d), dtype=np.float32)
    qs[:32] = pts[:32]                           # distance-0 hits with duplicates
    qs[32:64] = np.round(qs[32:64] * 8) / 8      # lattice queries against lattice points
    pts[100:400] = np.round(pts[100:400] * 8) / 8
    want = oracle.bf_knn(pts, qs, k, md)
    bf = knn.BruteForceGpu(dev(pts, cuda))
    for v in BF_VARIANTS:
        got = run(bf, qs, k, cuda, md, **v)
        w = want if not v["USE_HEAP_TRACKING"] else (want[0][:, ::-1].copy(), want[1][:, ::-1].copy())
        same(got, w, f"bf d={d} k={k} md={md} {v}")
    bf.close()
