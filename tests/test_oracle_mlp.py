"""The MLP oracle's gradient against central finite differences, and Adam
against its closed-form first step."""
import numpy as np

from oracle import mlp_ref as M


def test_gradient_finite_differences():
    rng = np.random.default_rng(0)
    nin, nh, nout, B = 6, 9, 2, 7
    p = rng.standard_normal(nin * nh + nh + nh * nout + nout)
    X = rng.standard_normal((B, nin))
    T = rng.standard_normal((B, nout))
    g = M.gradient(p, X, T, nin, nh, nout)
    eps = 1e-6
    fd = np.zeros_like(p)
    for i in range(len(p)):
        e = np.zeros_like(p)
        e[i] = eps
        fd[i] = (M.loss(p + e, X, T, nin, nh, nout) - M.loss(p - e, X, T, nin, nh, nout)) / (2 * eps)
    assert np.max(np.abs(g - fd)) <= 1e-7


def test_adam_first_step_is_signed_lr():
    p = np.array([1.0, -2.0, 3.0])
    g = np.array([0.5, -0.1, 0.0])
    p1, m, v = M.adam(p, g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    # bias-corrected first step: p - lr * g/|g| (0 where g == 0)
    assert np.allclose(p1, p - 0.01 * np.sign(g), atol=1e-6)
