"""GPU tests of the learned-radius form: MLP kernels against the FP64 oracle,
and tree results with the learned radius identical to the plain result."""
import ctypes

import numpy as np
import pytest

import cudaexactknn_amd as knn
from cudaexactknn_amd import _lib, datasets
from tests.test_gpu_parity import dev, run, same

pytestmark = pytest.mark.gpu


def test_mlp_gradient_matches_oracle(cuda):
    import torch
    from oracle import mlp
    g = np.random.Generator(np.random.Philox(11))
    p = (g.standard_normal(mlp.N_PARAMS) * 0.15).astype(np.float32)
    X = g.standard_normal((777, mlp.F_IN)).astype(np.float32)
    T = g.standard_normal(777).astype(np.float32)
    L = _lib.load()
    cnt = ctypes.c_uint32(0)
    _lib.check(L.knn_mlp_param_count(ctypes.byref(cnt)))
    assert cnt.value == mlp.N_PARAMS
    dp, dx, dt = dev(p, cuda), dev(X, cuda), dev(T, cuda)
    grad = torch.empty(mlp.N_PARAMS, dtype=torch.float32, device=cuda)
    loss = ctypes.c_float(0)
    _lib.check(L.knn_mlp_gradient(dp.data_ptr(), dx.data_ptr(), dt.data_ptr(), 777, grad.data_ptr(),
                                  ctypes.byref(loss)))
    want, want_loss = mlp.gradient(p.astype(np.float64), X, T)
    got = grad.cpu().numpy().astype(np.float64)
    scale = np.abs(want).max()
    err = np.abs(got - want).max() / scale
    print("max relative gradient error vs FP64 oracle:", err)
    assert err < 1e-4
    assert abs(loss.value - want_loss) <= 1e-4 * max(1.0, want_loss)
    out = torch.empty(777, dtype=torch.float32, device=cuda)
    _lib.check(L.knn_mlp_forward(dp.data_ptr(), dx.data_ptr(), 777, 0, out.data_ptr()))
    f32 = mlp.forward(p, X)
    assert np.max(np.abs(out.cpu().numpy() - f32)) <= 1e-4 * max(1.0, np.abs(f32).max())
    _lib.check(L.knn_mlp_forward(dp.data_ptr(), dx.data_ptr(), 777, 1, out.data_ptr()))
    b16 = mlp.forward(p, X, bf16_inference=True)
    assert np.max(np.abs(out.cpu().numpy() - b16)) <= 2e-3 * max(1.0, np.abs(b16).max())


@pytest.mark.parametrize("kind", ["kd", "ball"])
def test_learned_form_is_exact(oracle, cuda, kind):
    pts, qs, sample = datasets.anisotropic_mixture_8d(200000, 8000, 64, 8, seed=3, n_sample=4096)
    p = dev(pts, cuda)
    t = knn.KdTreeGpu(p, depth_skips=0) if kind == "kd" else knn.BallTreeGpu(p, depth_skips=1)
    import torch
    st = t.train_radius(torch.from_numpy(sample).to(cuda), 32, epochs=15)
    # The loss is the mean squared error of the standardised log2 K-th
    # radius^2: below 1 the network beats predicting the mean; here it must
    # explain most of the variance.
    print(kind, "training loss", st["loss"])
    assert np.isfinite(st["loss"]) and st["loss"] < 0.5
    want = oracle.bf_knn(pts, qs, 32)
    for radius in (False, True):
        for heap in (False, True):
            w = want if not heap else (want[0][:, ::-1].copy(), want[1][:, ::-1].copy())
            got = run(t, qs, 32, cuda, USE_HEAP_TRACKING=heap, form=knn.FORM_LEARNED, learned_radius=radius)
            same(got, w, f"learned {kind} heap={heap} radius={radius}")
        got = run(t, qs, 32, cuda, form=knn.FORM_LEARNED, learned_stats=True, learned_radius=radius)
        same(got, want, f"learned {kind} with statistics radius={radius}")
        s = t.last_query_stats()
        print(kind, "radius", radius, "short queries", s["rerun_queries"], "radius ratio", s["radius_ratio"],
              "far", s["far_fraction"])
        # 20 % of the queries are uniform over the bounding box, far from
        # the thin clusters: the predictor orders most of them as far, and
        # few mixture queries.
        assert 0.1 < s["far_fraction"] < 0.4
        if radius:
            assert s["rerun_queries"] < 0.05 * len(qs)
            assert s["radius_ratio"] >= 1.0 - 1e-6 or s["rerun_queries"] > 0
        else:
            assert s["rerun_queries"] == 0
        # The per-lane traversal takes the other path (with the radius, short
        # queries flagged and re-run by the plain traversal): same rows.
        same(run(t, qs, 32, cuda, form=knn.FORM_LEARNED, per_lane=True, learned_radius=radius), want,
             f"learned {kind} per-lane radius={radius}")


@pytest.mark.parametrize("kind", ["kd", "ball"])
def test_learned_second_pass_keeps_rows_exact(oracle, cuda, kind):
    """A predictor trained on queries that sit on the points themselves
    predicts radii far too small for queries spread over the bounding box:
    most of them end the first pass short and are finished by the in-kernel
    second pass (which keeps what the first pass found and accepts only
    points beyond the radius).  Rows stay exact, with finite and negative
    radii too."""
    import torch
    g = np.random.Generator(np.random.Philox(8))
    pts, _ = datasets.anisotropic_mixture_8d(60000, 10, 64, 8, seed=5)
    lo, hi = pts.min(0), pts.max(0)
    qs = (lo + (hi - lo) * g.random((3000, 8))).astype(np.float32)
    qs[:200] = pts[:200]
    p = dev(pts, cuda)
    t = knn.KdTreeGpu(p, depth_skips=0) if kind == "kd" else knn.BallTreeGpu(p, depth_skips=1)
    t.train_radius(torch.from_numpy(pts[1000:5096].copy()).to(cuda), 16, epochs=10)
    for md in (np.inf, 40.0, -1.0):
        for heap in (False, True):
            want = oracle.bf_knn(pts, qs, 16, max_d2=md, heap=heap)
            got = run(t, qs, 16, cuda, md=md, USE_HEAP_TRACKING=heap, form=knn.FORM_LEARNED, learned_radius=True)
            same(got, want, f"{kind} md={md} heap={heap}")
            if md == np.inf and not heap:
                assert t.last_query_stats()["rerun_queries"] > 100
