"""Scheduler MLP: the float64 oracle is pinned by finite differences (CPU);
the HIP forward / backward kernels are checked against it (GPU)."""
import numpy as np
import pytest

import vertexsurge_amd as vs
from oracle import mlp


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.random((n, mlp.IN)).astype(np.float32)
    Y = rng.normal(size=(n, mlp.OUT)).astype(np.float32)
    M = (rng.random((n, mlp.OUT)) < 0.5).astype(np.float32)
    M[:, 0] = 1
    return X, Y, M


def _params(seed):
    rng = np.random.default_rng(seed)
    return (rng.normal(size=mlp.NPARAM) * 0.2).astype(np.float64)


def test_oracle_gradient_matches_finite_differences():
    X, Y, M = _batch(16, 1)
    p = _params(2)
    L, g = mlp.loss_grad(p, X, Y, M)
    assert abs(L - mlp.loss(p, X, Y, M)) < 1e-12
    rng = np.random.default_rng(3)
    idx = rng.choice(mlp.NPARAM, size=60, replace=False)
    eps = 1e-6
    for i in idx:
        d = np.zeros_like(p)
        d[i] = eps
        fd = (mlp.loss(p + d, X, Y, M) - mlp.loss(p - d, X, Y, M)) / (2 * eps)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_oracle_catches_wrong_sign():
    X, Y, M = _batch(8, 4)
    p = _params(5)
    L, g = mlp.loss_grad(p, X, Y, M)
    # a step along +grad must increase the loss, along -grad decrease it
    assert mlp.loss(p + 1e-3 * g, X, Y, M) > L > mlp.loss(p - 1e-3 * g, X, Y, M)


@pytest.mark.gpu
def test_gpu_forward_and_gradient_match_oracle():
    s = vs.Scheduler(seed=7)
    X, Y, M = _batch(300, 8)
    p = s.get_params().astype(np.float64)
    y_ref, _ = mlp.forward(p, X)
    y = s.predict(X)
    assert np.allclose(y, y_ref, rtol=1e-4, atol=1e-5)
    L, g = s.loss_grad(X, Y, M)
    L_ref, g_ref = mlp.loss_grad(p, X, Y, M)
    assert abs(L - L_ref) <= 1e-4 * max(1.0, abs(L_ref))
    scale = np.abs(g_ref).max()
    assert np.allclose(g, g_ref, rtol=1e-3, atol=1e-5 * scale)
    # deterministic reduction: the same batch gives the same bits
    L2, g2 = s.loss_grad(X, Y, M)
    assert L2 == L and np.array_equal(g, g2)


@pytest.mark.gpu
def test_gpu_training_fits_a_known_function():
    s = vs.Scheduler(seed=3)
    rng = np.random.default_rng(0)
    X = rng.random((2000, mlp.IN)).astype(np.float32)
    Y = np.zeros((2000, mlp.OUT), dtype=np.float32)
    Y[:, 0] = 3 * X[:, 0] - 2 * X[:, 3] + 1
    Y[:, 1] = np.sin(3 * X[:, 1]) + X[:, 4]
    M = np.zeros_like(Y)
    M[:, :2] = 1
    curve = s.train(X, Y, M, epochs=600, lr=3e-3)
    assert curve[-1] < 0.05 * curve[0] and curve[-1] < 0.02
    p = s.get_params().astype(np.float64)
    assert abs(mlp.loss(p, X, Y, M) - s.loss_grad(X, Y, M)[0]) < 1e-4
