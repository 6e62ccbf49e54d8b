"""Pins the oracle MLP (CPU only): BF16 rounding on known values, GELU and
its derivative, and the backward pass against central differences of the
loss (with rounding switched off), including a planted slip that the check
must catch."""

import numpy as np

from oracle import mlp


def _net(seed, dims):
    rng = np.random.default_rng(seed)
    layers = [(rng.standard_normal((n, k)) / np.sqrt(k), 0.1 * rng.standard_normal(n))
              for k, n in zip(dims[:-1], dims[1:])]
    return layers, rng.standard_normal((9, dims[0])), rng.standard_normal((9, dims[-1]))


def _fd_grad(layers, dims, X, T):
    p = mlp.flatten(layers)
    fd = np.zeros_like(p)
    eps = 1e-6
    for i in range(p.size):
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd[i] = (mlp.loss_and_grad(mlp.unflatten(pp, dims), X, T, False)[0]
                 - mlp.loss_and_grad(mlp.unflatten(pm, dims), X, T, False)[0]) / (2 * eps)
    return fd


def test_bf16_rounding_known_values():
    x = np.array([1 + 2 ** -8, 1 + 3 * 2 ** -9, 1 + 2 ** -7 + 2 ** -9, -2.5, 3.14159, 0.0, 65504.0])
    assert np.array_equal(mlp.to_bf16(x), [1.0, 1.0078125, 1.0078125, -2.5, 3.140625, 0.0, 65536.0])


def test_gelu_values():
    assert mlp.gelu(0.0) == 0.0 and abs(mlp.gelu_grad(0.0) - 0.5) < 1e-15
    z = np.linspace(-4, 4, 41)
    fd = (mlp.gelu(z + 1e-6) - mlp.gelu(z - 1e-6)) / 2e-6
    assert np.allclose(fd, mlp.gelu_grad(z), atol=1e-8)
    assert abs(mlp.gelu(10.0) - 10.0) < 1e-12 and abs(mlp.gelu(-10.0)) < 1e-12


def test_backward_matches_central_differences():
    dims = [5, 8, 8, 3]
    layers, X, T = _net(0, dims)
    _, g = mlp.loss_and_grad(layers, X, T, bf16=False)
    gf = mlp.flatten(g)
    fd = _fd_grad(layers, dims, X, T)
    assert np.abs(fd - gf).max() <= 1e-7 * np.abs(gf).max()


def test_fd_check_catches_a_dropped_gelu_derivative(monkeypatch):
    dims = [5, 8, 8, 3]
    layers, X, T = _net(1, dims)
    monkeypatch.setattr(mlp, "gelu_grad", lambda z: np.ones_like(z))
    _, g = mlp.loss_and_grad(layers, X, T, bf16=False)
    monkeypatch.undo()
    fd = _fd_grad(layers, dims, X, T)
    assert np.abs(fd - mlp.flatten(g)).max() > 1e-2 * np.abs(fd).max()
