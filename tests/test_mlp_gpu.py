"""GPU kernels of the BFS variant predictor against the FP64 oracle
(oracle/mlp.py): inference, the gradient (against central finite
differences of the oracle's loss, independent of any backward pass), and
training (loss falls, a known decision rule is learnt, runs reproduce bit
for bit).  Tolerances: forward 1e-5 relative; gradient 2e-3 relative to the
gradient's norm (FP32 kernel vs FP64 finite differences)."""
import numpy as np
import pytest

from belewitte_amd import switch
from oracle import mlp as OM

pytestmark = pytest.mark.gpu
K = 3


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (n, 11)) * np.array([3, 4, 2, 5, 1, 1, 1, 2, 1, 2, 1]) \
        + np.array([10, 14, 20, 6, 3, 2, 1, 12, 20, -3, 0])
    return X.astype(np.float32)


def test_forward_matches_oracle(gpu):
    X = _data(1000, 1)
    m = switch.Model.init(X, K, seed=3, target_mean=12.0)
    got = m.predict(X).astype(np.float64)
    want = OM.forward(m.w, K, X.astype(np.float64))
    assert got.shape == (1000, K)
    assert np.max(np.abs(got - want)) <= 1e-5 * max(1.0, np.abs(want).max())


def test_gradient_matches_finite_differences(gpu):
    X = _data(100, 2)
    rng = np.random.default_rng(5)
    Y = rng.normal(12, 2, (100, K)).astype(np.float32)
    m = switch.Model.init(X, K, seed=4, target_mean=11.0)
    g, loss = m.gradient(X, Y)
    assert abs(loss - OM.loss(m.w, K, X, Y)) <= 1e-4 * OM.loss(m.w, K, X, Y)
    fd = OM.fd_gradient(m.w, K, X, Y, h=1e-4)
    assert np.all(g[:2 * OM.F] == 0)
    err = np.linalg.norm(g.astype(np.float64) - fd)
    assert err <= 2e-3 * np.linalg.norm(fd), (err, np.linalg.norm(fd))
    # Each block of parameters on its own (a wrong index inside one layer
    # would hide in the total norm).
    L = OM.layout(K)
    for a, b in (("W1", "b1"), ("b1", "W2"), ("W2", "b2"), ("b2", "W3"), ("W3", "b3"), ("b3", "end")):
        sl = slice(L[a], L[b])
        assert np.linalg.norm(g[sl] - fd[sl]) <= 5e-3 * max(np.linalg.norm(fd[sl]), 1e-3), a


def test_training_learns_a_rule(gpu):
    # Variant times with a crossover: variant 0 is fast on small frontiers,
    # variant 1 on large ones, variant 2 never best.
    X = _data(2000, 6)
    f = X[:, 0].astype(np.float64)
    Y = np.stack([f, 20.0 - 0.5 * f + 2.0, np.full_like(f, 30.0)], axis=1).astype(np.float32)
    m = switch.Model.init(X, K, seed=7, target_mean=float(Y.mean()))
    before = OM.loss(m.w, K, X, Y)
    losses = m.train(X, Y, epochs=400, batch=64, lr=3e-3, seed=11)
    after = OM.loss(m.w, K, X, Y)
    assert losses.shape == (400,) and np.all(np.isfinite(losses))
    assert after < 0.02 * before, (before, after)
    assert abs(losses[-1] - after) < 0.5 * after + 1e-3
    acc = np.mean(m.choose(X) == np.argmin(Y, axis=1))
    assert acc >= 0.97, acc


def test_training_is_reproducible(gpu):
    X = _data(300, 8)
    Y = np.random.default_rng(9).normal(10, 1, (300, K)).astype(np.float32)
    a = switch.Model.init(X, K, seed=1)
    b = switch.Model.init(X, K, seed=1)
    a.train(X, Y, epochs=50, batch=32, lr=1e-3, seed=2)
    b.train(X, Y, epochs=50, batch=32, lr=1e-3, seed=2)
    assert np.array_equal(a.w.view(np.uint32), b.w.view(np.uint32))
    assert np.array_equal(a.w[:2 * OM.F], switch.Model.init(X, K, seed=1).w[:2 * OM.F])  # mean/std untouched
