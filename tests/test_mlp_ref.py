"""Pins of the float64 network reference (tests/mlp_ref.py) against what the mathematics fixes: central finite
differences of the loss in every parameter, the closed form of the loss and gradient with zero weights, and Adam's
first step (a signed step of exactly lr in every coordinate with a nonzero gradient)."""
import numpy as np

import mlp_ref as R
from ampsolve import workloads as W


def params(nIn, H, nOut, seed):
    n = sum(int(np.prod(s)) for _, s in R.sizes(nIn, H, nOut))
    return 0.6 * (2 * W.lcg_uniform(seed, n) - 1)


def test_gradient_matches_finite_differences():
    nIn, H, nOut, B = 3, 5, 4, 6
    th = params(nIn, H, nOut, 11)
    X = 2 * W.lcg_uniform(12, B * nIn).reshape(B, nIn) - 1
    T = W.lcg_normal(13, B * nOut).reshape(B, nOut)
    _, g = R.loss_and_grad(th, X, T, nIn, H, nOut)
    fd = np.empty_like(th)
    h = 1e-6
    for i in range(len(th)):
        e = np.zeros_like(th)
        e[i] = h
        fd[i] = (R.loss_and_grad(th + e, X, T, nIn, H, nOut)[0] - R.loss_and_grad(th - e, X, T, nIn, H, nOut)[0]) / (2 * h)
    assert np.abs(fd - g).max() < 1e-8 * max(1.0, np.abs(g).max())


def test_zero_weights_closed_form():
    nIn, H, nOut, B = 2, 3, 4, 5
    th = np.zeros(sum(int(np.prod(s)) for _, s in R.sizes(nIn, H, nOut)))
    b3 = np.array([0.5, -1.0, 2.0, 0.0])
    th[-nOut:] = b3
    X = W.lcg_uniform(3, B * nIn).reshape(B, nIn)
    T = W.lcg_normal(4, B * nOut).reshape(B, nOut)
    loss, g = R.loss_and_grad(th, X, T, nIn, H, nOut)
    assert np.isclose(loss, np.mean((b3 - T) ** 2))
    assert np.allclose(g[-nOut:], 2 * (b3 - T).sum(0) / (B * nOut))
    assert np.all(g[:-nOut - H * nOut] == 0)  # tanh(0) = 0 blocks every earlier gradient through W3 = 0


def test_adam_first_step():
    g = np.array([0.3, -2.0, 0.0, 1e-3])
    th, m, v = R.adam(np.zeros(4), g, np.zeros(4), np.zeros(4), 1, 0.01)
    assert np.allclose(th[[0, 1, 3]], -0.01 * np.sign(g[[0, 1, 3]]), rtol=1e-4) and th[2] == 0
