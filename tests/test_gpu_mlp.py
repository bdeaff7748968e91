"""GPU kernels of the cost model's MLP (xc_mlp_*) against the float64 CPU reference
(tests/mlp_ref.py, itself pinned by finite differences): forward pass, loss and gradient,
Adam, and whole training runs (HIP-graph replay against eager launches)."""
import ctypes

import numpy as np
import pytest

import mlp_ref
from xc_testutil import xcorr

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _problem(dims, batch, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (batch, dims[0])).astype(np.float32)
    t = rng.normal(0, 1, (batch, dims[-1])).astype(np.float32)
    mask = (rng.random((batch, dims[-1])) > 0.1).astype(np.float32)
    return x, t, mask


def _init(x, dims, seed):
    sh = x.MlpShape.of(dims)
    p = np.zeros(x.lib().xc_mlp_num_params(ctypes.byref(sh)), np.float32)
    x.check(x.lib().xc_mlp_init(ctypes.byref(sh), seed, p.ctypes.data_as(ctypes.c_void_p)))
    return sh, p


@pytest.mark.parametrize("dims,batch", [([13, 32, 32, 8], 240), ([3, 5, 2], 7), ([13, 64, 8], 1000)])
def test_forward_and_loss_grad_match_cpu(dims, batch):
    import torch
    x = xcorr()
    sh, p = _init(x, dims, 11)
    assert p.size == mlp_ref.num_params(dims)
    # Glorot init: nonzero weights within the bound, zero biases
    W0, b0 = mlp_ref.split_params(dims, p.astype(np.float64))[0]
    assert np.abs(W0).max() <= np.sqrt(6.0 / (dims[0] + dims[1])) and np.abs(W0).max() > 0 and not b0.any()
    xs, ts, ms = _problem(dims, batch, 3)
    dp, dx, dt, dm = _dev(p), _dev(xs), _dev(ts), _dev(ms)
    work = torch.zeros(x.lib().xc_mlp_work_size(ctypes.byref(sh), batch), device="cuda")
    y = torch.zeros((batch, dims[-1]), device="cuda")
    x.check(x.lib().xc_mlp_forward_f32(ctypes.byref(sh), _p(dp), _p(dx), batch, _p(y), _p(work), None))
    loss = torch.zeros(1, device="cuda")
    grad = torch.zeros(p.size, device="cuda")
    x.check(x.lib().xc_mlp_loss_grad_f32(ctypes.byref(sh), _p(dp), _p(dx), _p(dt), _p(dm), batch, _p(loss),
                                         _p(grad), _p(work), None))
    torch.cuda.synchronize()
    y_ref, _ = mlp_ref.forward(dims, p, xs)
    np.testing.assert_allclose(y.cpu().numpy(), y_ref, rtol=1e-4, atol=1e-5)
    L_ref, g_ref = mlp_ref.loss_grad(dims, p, xs, ts, ms)
    assert loss.item() == pytest.approx(L_ref, rel=1e-5)
    g = grad.cpu().numpy()
    scale = np.abs(g_ref).max()
    assert np.abs(g - g_ref).max() <= 1e-4 * scale, (np.abs(g - g_ref).max(), scale)


def test_adam_matches_cpu():
    import torch
    x = xcorr()
    rng = np.random.default_rng(8)
    n = 1000
    p = rng.normal(size=n).astype(np.float32)
    pm, pv = np.zeros(n), np.zeros(n)
    dp, dm, dv = _dev(p), torch.zeros(n, device="cuda"), torch.zeros(n, device="cuda")
    pr = p.astype(np.float64)
    for step in range(1, 6):
        g = rng.normal(size=n).astype(np.float32)
        dg = _dev(g)
        x.check(x.lib().xc_mlp_adam_f32(n, _p(dp), _p(dg), _p(dm), _p(dv), step, 1e-2, 0.9, 0.999, 1e-8, 1e-3, None))
        pr, pm, pv = mlp_ref.adam(pr, g.astype(np.float64), pm, pv, step, 1e-2, wd=1e-3)
    torch.cuda.synchronize()
    np.testing.assert_allclose(dp.cpu().numpy(), pr, rtol=1e-5, atol=1e-6)
    assert x.lib().xc_mlp_adam_f32(n, _p(dp), _p(dg), _p(dm), _p(dv), 0, 1e-2, 0.9, 0.999, 1e-8, 0.0, None) == 1


def _train(x, sh, p0, xs, ts, ms, epochs, use_graph, lr=1e-2, wd=0.0):
    p = p0.copy()
    loss = np.zeros(epochs, np.float32)
    o = x.TrainOptions(epochs, lr, wd, 0, 0, int(use_graph))
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    x.check(x.lib().xc_mlp_train_f32(ctypes.byref(sh), vp(xs), vp(ts), vp(ms), xs.shape[0], ctypes.byref(o), vp(p),
                                     vp(loss)))
    return p, loss


def test_training_graph_equals_eager_and_cpu():
    x = xcorr()
    dims = [13, 16, 16, 8]
    sh, p0 = _init(x, dims, 5)
    xs, ts, ms = _problem(dims, 64, 9)
    # Reproducible, and the graph replays (each on the same device step counter) give the
    # same parameters and loss curve as eager launches, bit for bit.
    pg, lg = _train(x, sh, p0, xs, ts, ms, 40, True)
    pe, le = _train(x, sh, p0, xs, ts, ms, 40, False)
    pg2, lg2 = _train(x, sh, p0, xs, ts, ms, 40, True)
    assert np.array_equal(pg, pe) and np.array_equal(lg, le)
    assert np.array_equal(pg, pg2) and np.array_equal(lg, lg2)
    # Same run on the CPU reference.
    p = p0.astype(np.float64)
    m, v = np.zeros_like(p), np.zeros_like(p)
    ref_loss = []
    for step in range(1, 41):
        L, g = mlp_ref.loss_grad(dims, p, xs, ts, ms)
        ref_loss.append(L)
        p, m, v = mlp_ref.adam(p, g, m, v, step, 1e-2)
    np.testing.assert_allclose(lg, ref_loss, rtol=2e-4)
    np.testing.assert_allclose(pg, p, rtol=2e-3, atol=2e-4)


def test_training_fits_a_smooth_target():
    x = xcorr()
    dims = [4, 32, 32, 2]
    sh, p0 = _init(x, dims, 2)
    rng = np.random.default_rng(4)
    xs = rng.uniform(-1, 1, (256, 4)).astype(np.float32)
    ts = np.stack([np.sin(xs[:, 0]) + xs[:, 1] * xs[:, 2], np.cos(xs[:, 3]) - xs[:, 0] ** 2], 1).astype(np.float32)
    ms = np.ones_like(ts)
    p, loss = _train(x, sh, p0, xs, ts, ms, 1500, True, lr=5e-3)
    assert loss[-1] < 0.02 * loss[0], (loss[0], loss[-1])
    # the trained parameters give a small loss on the CPU reference as well
    assert mlp_ref.loss(dims, p.astype(np.float64), xs, ts, ms) < 0.02 * loss[0]
