"""Pins of the numpy restatement of the warm-start network: the reverse-mode gradient against
central finite differences of the loss, and the forward pass against a hand-written two-layer
evaluation."""
import numpy as np

import mlp as M


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(0)
    widths = [5, 7, 6, 3]
    p = rng.normal(0, 0.5, M.n_params(widths))
    X = rng.normal(size=(11, 5))
    T = rng.normal(size=(11, 3))
    loss, g = M.loss_and_gradient(p, widths, X, T)
    h = 1e-6
    for k in rng.choice(p.size, 40, replace=False):
        e = np.zeros_like(p)
        e[k] = h
        fd = (M.loss_and_gradient(p + e, widths, X, T)[0] - M.loss_and_gradient(p - e, widths, X, T)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_forward_is_the_stated_function():
    rng = np.random.default_rng(1)
    widths = [3, 4, 2]
    p = rng.normal(size=M.n_params(widths))
    W1 = p[:12].reshape(4, 3); b1 = p[12:16]; W2 = p[16:24].reshape(2, 4); b2 = p[24:26]
    x = rng.normal(size=(1, 3))
    y = np.tanh(x @ W1.T + b1) @ W2.T + b2
    np.testing.assert_allclose(M.forward(p, widths, x)[-1], y, rtol=1e-14)
