"""Seeded synthetic point sets for the five workloads the library is measured on.

Every generator draws from numpy's counter-based Philox bit generator, so a
(seed, size) pair gives the same float32 arrays on every machine.  Sizes can
be scaled down for tests while keeping the recipe.
"""
from __future__ import annotations

import numpy as np


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(seed))


def uniform_cube(n_points: int, n_queries: int, dims: int = 3, seed: int = 0):
    """Points and queries i.i.d. uniform in [0,1)^dims (small CPU-runnable case:
    65,536 points, 16,384 queries, D=3, seed 0)."""
    g = _rng(seed)
    pts = g.random((n_points, dims), dtype=np.float32)
    qs = g.random((n_queries, dims), dtype=np.float32)
    return pts, qs


def gaussian_mixture_3d(n_points: int = 1 << 24, n_queries: int = 1 << 22, n_clusters: int = 32,
                        seed: int = 1):
    """Mixture of isotropic Gaussians in R^3: centres uniform in [0,100)^3, sigma
    log-uniform in [0.5, 5], equal weights; queries from the same mixture."""
    g = _rng(seed)
    centres = g.uniform(0.0, 100.0, (n_clusters, 3))
    sigma = np.exp(g.uniform(np.log(0.5), np.log(5.0), n_clusters))

    def draw(m):
        lab = g.integers(0, n_clusters, m)
        x = centres[lab] + g.standard_normal((m, 3)) * sigma[lab, None]
        return x.astype(np.float32)

    return draw(n_points), draw(n_queries)


def scaled_normal(n_points: int = 1 << 20, n_queries: int = 1 << 16, dims: int = 128, seed: int = 2):
    """N(0,1) per coordinate, then each dimension scaled by a factor log-uniform
    in [0.1, 10] shared by points and queries."""
    g = _rng(seed)
    scale = np.exp(g.uniform(np.log(0.1), np.log(10.0), dims))
    pts = (g.standard_normal((n_points, dims)) * scale).astype(np.float32)
    qs = (g.standard_normal((n_queries, dims)) * scale).astype(np.float32)
    return pts, qs


def anisotropic_mixture_8d(n_points: int = 1 << 22, n_queries: int = 1 << 20, n_clusters: int = 64,
                           dims: int = 8, seed: int = 3, uniform_fraction: float = 0.2, n_sample: int = 0):
    """64 rotated anisotropic Gaussians in R^8 (centres uniform in [0,10)^8, axis
    scales log-uniform in [0.01, 1], weights Dirichlet(1)); 80 % of queries from
    the mixture, 20 % uniform over the points' bounding box.  With n_sample > 0
    a third array follows: n_sample more queries from the same mixture and
    recipe (drawn after the others, which stay as they are), the learned
    radius's training sample."""
    g = _rng(seed)
    centres = g.uniform(0.0, 10.0, (n_clusters, dims))
    scales = np.exp(g.uniform(np.log(0.01), np.log(1.0), (n_clusters, dims)))
    rots = []
    for _ in range(n_clusters):
        a = g.standard_normal((dims, dims))
        qm, r = np.linalg.qr(a)
        qm = qm * np.sign(np.diag(r))[None, :]
        rots.append(qm)
    rots = np.stack(rots)
    weights = g.dirichlet(np.ones(n_clusters))

    def draw(m):
        lab = g.choice(n_clusters, size=m, p=weights)
        z = g.standard_normal((m, dims)) * scales[lab]
        x = centres[lab] + np.einsum("mij,mj->mi", rots[lab], z)
        return x.astype(np.float32)

    pts = draw(n_points)
    lo, hi = pts.min(axis=0).astype(np.float64), pts.max(axis=0).astype(np.float64)

    def queries(m):
        n_uni = int(round(m * uniform_fraction))
        q_mix = draw(m - n_uni)
        q_uni = (lo + (hi - lo) * g.random((n_uni, dims))).astype(np.float32)
        q = np.concatenate([q_mix, q_uni])
        return q[g.permutation(m)]

    qs = queries(n_queries)
    if n_sample:
        return pts, qs, queries(n_sample)
    return pts, qs


def uneven_batch(n_sets: int = 256, lo: int = 1024, hi: int = 1 << 20, seed: int = 4):
    """Independent datasets: sizes log-uniform in [lo, hi], D uniform over {3,16},
    uniform-in-unit-cube points, queries = size/4, K uniform over {4,8,16}."""
    g = _rng(seed)
    sets = []
    for _ in range(n_sets):
        n = int(np.exp(g.uniform(np.log(lo), np.log(hi))))
        d = int(g.choice([3, 16]))
        k = int(g.choice([4, 8, 16]))
        pts = g.random((n, d), dtype=np.float32)
        qs = g.random((max(1, n // 4), d), dtype=np.float32)
        sets.append({"points": pts, "queries": qs, "dims": d, "k": k})
    return sets
