"""HIP MLP kernels (forward, backward, Adam) against the NumPy oracle."""
import numpy as np
import pytest

from oracle import mlp_ref as M

pytestmark = pytest.mark.gpu


def _data(B=37, nin=66, nout=2, seed=3):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((B, nin)).astype(np.float32),
            rng.standard_normal((B, nout)).astype(np.float32))


def test_forward_and_gradient_match_oracle(ctx):
    from sparse_matrix_algorithms_amd.learned import MLP
    m = MLP(66, 128, 2, seed=9, ctx=ctx)
    X, T = _data()
    p = m.params().astype(np.float64)
    Y = m.predict(X)
    Yo, _ = M.forward(p, X.astype(np.float64), 66, 128, 2)
    assert np.max(np.abs(Y - Yo)) <= 1e-4 * max(1.0, np.abs(Yo).max())
    loss, g = m.gradient(X, T)
    go = M.gradient(p, X.astype(np.float64), T.astype(np.float64), 66, 128, 2)
    assert abs(loss - M.loss(p, X.astype(np.float64), T.astype(np.float64), 66, 128, 2)) <= 1e-4
    assert np.max(np.abs(g - go)) <= 1e-4 * max(1.0, np.abs(go).max())


def test_adam_steps_match_oracle(ctx):
    from sparse_matrix_algorithms_amd.learned import MLP
    m = MLP(66, 128, 2, seed=4, ctx=ctx)
    X, T = _data(B=64)
    p = m.params().astype(np.float64)
    mo, vo = np.zeros_like(p), np.zeros_like(p)
    for t in range(1, 6):
        g = M.gradient(p, X.astype(np.float64), T.astype(np.float64), 66, 128, 2)
        p, mo, vo = M.adam(p, g, mo, vo, t, lr=1e-3)
    m.train(X, T, 5, 1e-3)
    assert np.max(np.abs(m.params() - p)) <= 1e-4


def test_training_reduces_loss(ctx):
    from sparse_matrix_algorithms_amd.learned import MLP
    rng = np.random.default_rng(1)
    X = rng.standard_normal((128, 66)).astype(np.float32)
    W = rng.standard_normal((66, 2)).astype(np.float32) / 8
    T = np.tanh(X @ W).astype(np.float32)
    m = MLP(66, 128, 2, seed=2, ctx=ctx)
    hist = m.train(X, T, 400, 3e-3)
    assert hist[-1] < 0.1 * hist[0]
