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
    pts, qs = datasets.anisotropic_mixture_8d(200000, 12000, 64, 8, seed=3)
    p = dev(pts, cuda)
    t = knn.KdTreeGpu(p, depth_skips=0) if kind == "kd" else knn.BallTreeGpu(p, depth_skips=1)
    import torch
    sample = torch.from_numpy(qs[:4096]).to(cuda)
    st = t.train_radius(sample, 32, epochs=15)
    assert np.isfinite(st["loss"]) and st["loss"] < 1.0
    test_q = qs[4096:]
    want = oracle.bf_knn(pts, test_q, 32)
    for heap in (False, True):
        w = want if not heap else (want[0][:, ::-1].copy(), want[1][:, ::-1].copy())
        got = run(t, test_q, 32, cuda, USE_HEAP_TRACKING=heap, form=knn.FORM_LEARNED)
        same(got, w, f"learned {kind} heap={heap}")
    s = t.last_query_stats()
    print(kind, "reruns", s["rerun_queries"], "radius ratio", s["radius_ratio"])
    assert s["rerun_queries"] < 0.5 * len(test_q)
    # A radius of 0 for everyone: every query re-runs and results stay exact.
    got = run(t, test_q, 32, cuda, md=np.inf, form=knn.FORM_LEARNED)
    same(got, want, "learned again")
