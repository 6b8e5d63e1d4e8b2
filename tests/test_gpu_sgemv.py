"""GPU parity tests of the SGEMV path (hmlSgemv / hmlSgemvEx).

Oracles as for SGEMM: the restated reference CPU path
(apps/gemv/hml_sgemv_main.cu:254-283), FP64, and integer-valued inputs that
make the result exact (bit-exact comparison) in any summation order.
"""
import ctypes

import numpy as np
import pytest

from util import assert_within_gamma, f64_ref, gamma, int_matrix, uniform

pytestmark = pytest.mark.gpu


def run_gemv(hml, A, x, y0, alpha, beta, order="row", trans=False,
             variant=None, flags=0):
    M, N = A.shape
    As = np.ascontiguousarray(A if order == "row" else A.T)
    lda = N if order == "row" else M
    dA = hml.DeviceBuffer.from_numpy(As)
    dx = hml.DeviceBuffer.from_numpy(x)
    dy = hml.DeviceBuffer.from_numpy(y0)
    hml.sgemv_ex(M, N, dA, dx, dy, alpha, beta, order=order, trans=trans,
                 lda=lda, variant=variant, flags=flags)
    hml.synchronize()
    return dy.to_numpy(y0.shape)


def exact(A, x, trans, alpha=1.0, beta=0.0, y0=None):
    Ai = A.astype(np.int64)
    r = (Ai.T if trans else Ai) @ x.astype(np.int64)
    r = alpha * r.astype(np.float64)
    if y0 is not None and beta != 0.0:
        r = r + beta * y0.astype(np.float64)
    return r.astype(np.float32)


SHAPES = [(1000, 64), (64, 1000), (4097, 129), (1, 300), (300, 1), (17, 17),
          (20000, 3), (5, 70000), (2048, 2048), (1 << 20, 64)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "%dx%d" % s)
@pytest.mark.parametrize("order", ["row", "col"])
@pytest.mark.parametrize("trans", [False, True], ids=["N", "T"])
def test_integer_exact(hml, shape, order, trans):
    M, N = shape
    print("case", shape, order, trans, flush=True)
    rng = np.random.default_rng(M * 7 + N)
    A = int_matrix(rng, (M, N))
    nx, ny = (M, N) if trans else (N, M)
    x = int_matrix(rng, (nx,))
    y0 = int_matrix(rng, (ny,))
    for alpha, beta in ((1.0, 0.0), (3.0, -2.0)):
        y = run_gemv(hml, A, x, y0, alpha, beta, order, trans)
        np.testing.assert_array_equal(y, exact(A, x, trans, alpha, beta, y0))


CASES = [(16384, 64, "row", False), (64, 16384, "row", False),
         (16384, 64, "row", True), (3000, 700, "row", True),
         (3000, 700, "col", False), (100, 100000, "row", False)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%dx%d-%s-%s" % (c[0], c[1], c[2], "T" if c[3] else "N"))
def test_integer_exact_every_candidate(hml, case):
    M, N, order, trans = case
    cands = (hml.SgemvVariant * 512)()
    n = hml.lib.hmlSgemvCandidates(0 if order == "row" else 1, int(trans), M, N, cands, 512)
    assert 0 < n <= 512
    rng = np.random.default_rng(5)
    A = int_matrix(rng, (M, N))
    nx, ny = (M, N) if trans else (N, M)
    x = int_matrix(rng, (nx,))
    y0 = int_matrix(rng, (ny,))
    R = exact(A, x, trans, 1.0, 1.0, y0)
    for i in range(n):
        print("variant", cands[i], flush=True)
        y = run_gemv(hml, A, x, y0, 1.0, 1.0, order, trans, variant=cands[i])
        np.testing.assert_array_equal(y, R, err_msg=repr(cands[i]))


def test_reference_inputs_compat(hml):
    """hmlSgemv with the reference app's inputs (rand()/RAND_MAX) against
    the restated reference CPU path and FP64."""
    import oracle
    for M, N in ((1000, 1000), (4096, 64), (50, 5000)):
        A, x, y = oracle.hml_gemv_inputs(M, N)
        yg = y.copy()
        hml.sgemv(yg, A, x, 1.0, 0.0)
        yc = oracle.sgemv_cpu(y.copy(), A, x, 1.0, 0.0)
        R, bound = f64_ref(A, x[:, None])
        assert_within_gamma(yg[:, None], R, bound, N)
        assert_within_gamma(yc[:, None], R, bound, N)
        rel = abs(oracle.ysum(yg) - oracle.ysum(yc)) / oracle.ysum(yc)
        print("ysum relative difference: %.20f" % rel)
        assert rel < 1e-6
        yb = y.copy()
        hml.sgemv(yb, A, x, 1.0, 0.0, basic=True)
        assert_within_gamma(yb[:, None], R, bound, N)


def test_random_vs_fp64_layouts(hml):
    rng = np.random.default_rng(2)
    for (M, N) in ((513, 257), (31, 4099), (70000, 16)):
        for order in ("row", "col"):
            for trans in (False, True):
                A = uniform(rng, (M, N))
                nx, ny = (M, N) if trans else (N, M)
                x = uniform(rng, (nx,))
                y0 = uniform(rng, (ny,))
                y = run_gemv(hml, A, x, y0, -1.25, 0.5, order, trans)
                At = A.T if trans else A
                R, bound = f64_ref(At, x[:, None], -1.25, 0.5, y0[:, None])
                assert_within_gamma(y[:, None], R, bound, nx)


def test_compat_staged_copies(hml, monkeypatch):
    """hmlSgemv with a 92 MiB host A (three pinned 32 MiB chunks, both
    staging buffers reused): integer-valued inputs give the exact result,
    with staging on and with HML_HOST_STAGING=off."""
    rng = np.random.default_rng(12)
    M, N = 6000, 4000
    A = int_matrix(rng, (M, N))
    x = int_matrix(rng, (N,))
    y0 = int_matrix(rng, (M,))
    R = (A.astype(np.int64) @ x.astype(np.int64) + 3 * y0.astype(np.int64)).astype(np.float32)
    for mode in ("on", "off"):
        monkeypatch.setenv("HML_HOST_STAGING", mode)
        y = y0.copy()
        hml.sgemv(y, A, x, 1.0, 3.0)
        np.testing.assert_array_equal(y, R, err_msg="staging " + mode)


def test_compat_reads_y_when_beta_zero(hml):
    rng = np.random.default_rng(4)
    A = uniform(rng, (64, 40))
    x = uniform(rng, (40,))
    y0 = np.zeros(64, np.float32)
    y0[5] = np.nan
    yc = y0.copy()
    hml.sgemv(yc, A, x, 1.0, 0.0)
    assert np.isnan(yc[5]) and np.isfinite(np.delete(yc, 5)).all()
    yb = run_gemv(hml, A, x, y0, 1.0, 0.0)
    assert np.isfinite(yb).all()


@pytest.mark.parametrize("trans", [False, True], ids=["N", "T"])
@pytest.mark.parametrize("shape", [(16384, 16384), (1 << 20, 64)], ids=["16384sq", "1Mx64"])
def test_full_size_vs_fp64(hml, shape, trans):
    M, N = shape
    print("case", shape, trans, flush=True)
    nx, ny = (M, N) if trans else (N, M)
    dA = hml.DeviceBuffer(4 * M * N)
    dx = hml.DeviceBuffer(4 * nx)
    dy = hml.DeviceBuffer(4 * ny)
    dA.fill_uniform(M * N, 201)
    dx.fill_uniform(nx, 202)
    hml.sgemv_ex(M, N, dA, dx, dy, 1.0, 0.0, trans=trans)
    # y = op(A) x is the GEMM (ny x 1) = op(A) (ny x nx) * x (nx x 1)
    st = hml.check_f64(ny, 1, nx, dA, dx, dy, transA=trans, lda=N, ldb=1, ldc=1)
    print("rel_fro %.3e bound %.3e" % (st["rel_fro"], gamma(nx) * st["bound_fro"]))
    assert st["nonfinite"] == 0
    assert st["max_rel_bound"] <= 2 * gamma(nx + 2)


@pytest.mark.parametrize("trans", [False, True], ids=["N", "T"])
def test_unaligned_views_every_candidate(hml, trans):
    """A starts one float past a 16-byte boundary and has an odd leading
    dimension (NaN padding, so any read of it shows up), x starts one float
    in: the buffer-load kernels for dword-aligned rows.  Ragged tails (N and
    M not multiples of 4), integer data, bit-exact for every candidate."""
    rng = np.random.default_rng(11)
    for (M, N, pad) in ((1003, 517, 3), (70, 2049, 1), (4099, 61, 5)):
        lda = N + pad
        A = int_matrix(rng, (M, N))
        Ap = np.full((M, lda), np.nan, np.float32)
        Ap[:, :N] = A
        nx, ny = (M, N) if trans else (N, M)
        x = int_matrix(rng, (nx,))
        y0 = int_matrix(rng, (ny,))
        dA = hml.DeviceBuffer.from_numpy(np.concatenate([[np.float32(7)], Ap.ravel()]))
        dx = hml.DeviceBuffer.from_numpy(np.concatenate([[np.float32(5)], x]).astype(np.float32))
        R = exact(A, x, trans, 1.0, 1.0, y0)
        cands = (hml.SgemvVariant * 512)()
        n = hml.lib.hmlSgemvCandidates(0, int(trans), M, N, cands, 512)
        assert 0 < n <= 512
        for i in [None] + list(range(n)):
            v = None if i is None else cands[i]
            print("case", M, N, pad, "T" if trans else "N", v, flush=True)
            dy = hml.DeviceBuffer.from_numpy(y0)
            hml.sgemv_ex(M, N, dA.ptr + 4, dx.ptr + 4, dy, 1.0, 1.0, trans=trans,
                         lda=lda, variant=v)
            hml.synchronize()
            np.testing.assert_array_equal(dy.to_numpy(y0.shape), R, err_msg=repr(v))
