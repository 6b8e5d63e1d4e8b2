"""Pins the numpy MLP oracle (oracle/mlp_oracle.py) that the GPU training
kernels are checked against: analytic gradients equal central finite
differences of the oracle's own loss; closed forms of the loss."""
import numpy as np

import mlp_oracle as mo


def problem(seed=0, B=7, dims=(6, 5, 4, 9)):
    rng = np.random.default_rng(seed)
    n = sum(dims[i - 1] * dims[i] + dims[i] for i in range(1, len(dims)))
    params = rng.normal(0, 0.7, n)
    X = rng.normal(0, 1, (B, dims[0]))
    mask = (rng.random((B, dims[-1])) < 0.7).astype(np.float64)
    mask[:, 0] = 1.0
    T = rng.random((B, dims[-1])) * mask
    T /= T.sum(axis=1, keepdims=True)
    return list(dims), params, X, T, mask


def test_gradient_matches_finite_differences():
    dims, params, X, T, mask = problem()
    L, g = mo.loss_grad(params, dims, X, T, mask)
    assert abs(L - mo.loss(params, dims, X, T, mask)) < 1e-12
    h = 1e-6
    fd = np.empty_like(params)
    for i in range(len(params)):
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        fd[i] = (mo.loss(p1, dims, X, T, mask) - mo.loss(p2, dims, X, T, mask)) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-8)


def test_closed_forms():
    dims = [6, 4, 4, 5]
    n = sum(dims[i - 1] * dims[i] + dims[i] for i in range(1, len(dims)))
    X = np.ones((3, 6))
    mask = np.array([[1, 1, 1, 1, 1], [1, 0, 1, 0, 0], [0, 0, 0, 1, 0]], float)
    T = mask / mask.sum(axis=1, keepdims=True)
    P = mo.masked_softmax(mo.forward(np.zeros(n), dims, X)[-1], mask)
    np.testing.assert_allclose(P, T)                       # zero net: uniform over the mask
    onehot = np.zeros((3, 5)); onehot[0, 2] = onehot[1, 0] = onehot[2, 3] = 1
    want = -(np.log(1 / 5) + np.log(1 / 2) + np.log(1.0)) / 3
    assert abs(mo.loss(np.zeros(n), dims, X, onehot, mask) - want) < 1e-12
