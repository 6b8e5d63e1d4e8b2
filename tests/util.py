"""Shared helpers of the test suite (layouts, seeded inputs, error bounds)."""
import numpy as np

U32 = 2.0 ** -24


def gamma(K, u=U32):
    """Higham's gamma_K = K u / (1 - K u)."""
    return K * u / (1.0 - K * u)


def store(X, order, trans):
    """Storage of op(X) == X for a BLAS call with (order, trans).

    Returns (contiguous array to upload, leading dimension).  The stored
    matrix is X itself in row-major/N and column-major/T, and X^T otherwise.
    """
    rows, cols = X.shape
    if (order == "row") == (not trans):
        return np.ascontiguousarray(X), max(1, cols)
    return np.ascontiguousarray(X.T), max(1, rows)


def store_c(C, order):
    if order == "row":
        return np.ascontiguousarray(C), max(1, C.shape[1])
    return np.ascontiguousarray(C.T), max(1, C.shape[0])


def load_c(buf, M, N, order):
    if order == "row":
        return buf.to_numpy((M, N))
    return buf.to_numpy((N, M)).T.copy()


def int_matrix(rng, shape, lo=-2, hi=2):
    return rng.integers(lo, hi + 1, size=shape).astype(np.float32)


def uniform(rng, shape, lo=-1.0, hi=1.0):
    return rng.uniform(lo, hi, size=shape).astype(np.float32)


def loguniform_signed(rng, shape, lo_exp=-3.0, hi_exp=3.0):
    mag = 10.0 ** rng.uniform(lo_exp, hi_exp, size=shape)
    sign = np.where(rng.random(size=shape) < 0.5, -1.0, 1.0)
    return (mag * sign).astype(np.float32)


def f64_ref(A, B, alpha=1.0, beta=0.0, C0=None):
    """FP64 reference and the componentwise bound term |alpha||A||B| +
    |beta||C0| (numpy float64 matmul, an independent library routine)."""
    A64 = A.astype(np.float64)
    B64 = B.astype(np.float64)
    R = alpha * (A64 @ B64)
    bound = abs(alpha) * (np.abs(A64) @ np.abs(B64))
    if C0 is not None and beta != 0.0:
        R = R + beta * C0.astype(np.float64)
        bound = bound + abs(beta) * np.abs(C0.astype(np.float64))
    return R, bound


def assert_within_gamma(C, R, bound, K, slack=2.0):
    """Componentwise |C - R| <= slack * gamma_{K+2} * bound (Higham 3.5)."""
    err = np.abs(C.astype(np.float64) - R)
    lim = slack * gamma(K + 2) * bound + 1e-30
    bad = err > lim
    assert not bad.any(), (
        "max err %.3e exceeds gamma bound at %d entries (first %s)"
        % (err.max(), int(bad.sum()), np.argwhere(bad)[:3].tolist()))
