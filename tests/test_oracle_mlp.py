"""Pins the FP64 MLP restatement (oracle/mlp.py) before it checks the GPU
kernels: with hand-built weights the network reduces to a closed form, and
the finite-difference gradient of the output biases equals the analytic
mean-squared-error derivative."""
import numpy as np

from oracle import mlp as OM

K = 3


def _identity_net(shift=(1.0, 2.0, 3.0), mean=0.0, std=1.0):
    L = OM.layout(K)
    w = np.zeros(L["end"])
    w[L["mean"]:L["std"]] = mean
    w[L["std"]:L["W1"]] = std
    W1 = np.zeros((32, OM.F))
    W1[:OM.F, :OM.F] = np.eye(OM.F)
    w[L["W1"]:L["b1"]] = W1.ravel()
    w[L["W2"]:L["b2"]] = np.eye(32).ravel()
    W3 = np.zeros((K, 32))
    W3[np.arange(K), np.arange(K)] = 1.0
    w[L["W3"]:L["b3"]] = W3.ravel()
    w[L["b3"]:L["end"]] = shift
    return w


def test_layout_sizes():
    assert OM.F == 11
    assert OM.layout(3)["end"] == 22 + 352 + 32 + 1024 + 32 + 96 + 3
    assert OM.layout(16)["end"] == 22 + 352 + 32 + 1024 + 32 + 512 + 16


def test_forward_closed_form():
    X = np.array([[-1.0, 2.0, 0.5] + [7.0] * (OM.F - 3), [3.0, -4.0, -0.25] + [0.0] * (OM.F - 3)])
    out = OM.forward(_identity_net(), K, X)
    want = np.maximum(X[:, :K], 0) + np.array([1.0, 2.0, 3.0])
    assert np.allclose(out, want, atol=0, rtol=0)
    # normalisation: (x - mean) / std before the first layer
    out2 = OM.forward(_identity_net(mean=1.0, std=2.0), K, X)
    assert np.allclose(out2, np.maximum((X[:, :K] - 1.0) / 2.0, 0) + [1.0, 2.0, 3.0])


def test_fd_gradient_of_output_bias():
    rng = np.random.default_rng(0)
    X = rng.normal(0, 2, (20, OM.F))
    Y = rng.normal(0, 1, (20, K))
    w = _identity_net()
    L = OM.layout(K)
    idx = list(range(L["b3"], L["end"]))
    g = OM.fd_gradient(w, K, X, Y, h=1e-5, indices=idx)
    r = OM.forward(w, K, X) - Y
    want = 2.0 * r.sum(axis=0) / r.size
    assert np.allclose(g[idx], want, rtol=1e-7, atol=1e-9)
    assert np.all(g[:L["b3"]] == 0)
