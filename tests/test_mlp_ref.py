"""Pins the CPU MLP reference (tests/mlp_ref.py) that checks the GPU kernels: its gradient
against central finite differences, its forward pass against a hand-computed case, and
Adam's first step against its closed form."""
import numpy as np
import pytest

import mlp_ref


@pytest.mark.parametrize("dims", [[3, 4, 2], [5, 7, 6, 3], [13, 8, 8, 8]])
def test_gradient_matches_finite_differences(dims):
    rng = np.random.default_rng(sum(dims))
    p = rng.normal(0, 0.5, mlp_ref.num_params(dims))
    x = rng.normal(0, 1, (6, dims[0]))
    t = rng.normal(0, 1, (6, dims[-1]))
    mask = (rng.random((6, dims[-1])) > 0.2).astype(np.float64)
    L, g = mlp_ref.loss_grad(dims, p, x, t, mask)
    assert L == pytest.approx(mlp_ref.loss(dims, p, x, t, mask), rel=1e-12)
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (mlp_ref.loss(dims, p + e, x, t, mask) - mlp_ref.loss(dims, p - e, x, t, mask)) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-8)


def test_forward_hand_case():
    # 1 input -> 1 hidden (tanh) -> 1 output: y = w2 * tanh(w1 * x + b1) + b2
    dims = [1, 1, 1]
    p = np.array([0.5, 0.1, 2.0, -0.3])
    y, _ = mlp_ref.forward(dims, p, np.array([[2.0]]))
    assert y[0, 0] == pytest.approx(2.0 * np.tanh(1.1) - 0.3, rel=1e-14)


def test_masked_entries_do_not_contribute():
    dims = [2, 3, 2]
    rng = np.random.default_rng(5)
    p = rng.normal(size=mlp_ref.num_params(dims))
    x = rng.normal(size=(4, 2))
    t = rng.normal(size=(4, 2))
    mask = np.ones((4, 2))
    mask[1, 0] = 0
    t2 = t.copy()
    t2[1, 0] += 100.0
    assert mlp_ref.loss_grad(dims, p, x, t, mask)[0] == pytest.approx(mlp_ref.loss_grad(dims, p, x, t2, mask)[0])


def test_adam_first_step_is_signed_lr():
    # At step 1 the bias-corrected update is lr * g / (|g| + eps) = lr * sign(g).
    g = np.array([0.3, -2.0, 1e-3])
    p, m, v = mlp_ref.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, 0.01, eps=0.0)
    np.testing.assert_allclose(p, -0.01 * np.sign(g), rtol=1e-12)
