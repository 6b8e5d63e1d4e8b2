"""hmlSgemmExNoAlloc: the SGEMM entry point that allocates nothing (partial
sums in a fixed device scratch array) against hmlSgemmEx, called through the
C ABI.  Both pick the same variant, so results must agree bit for bit; with
integer-valued inputs both must also equal the exact product."""
import ctypes

import numpy as np
import pytest

from util import int_matrix

pytestmark = pytest.mark.gpu

SCRATCH = 16 << 20  # HML_NOALLOC_SCRATCH_BYTES


def _call(hml, fn, M, N, K, dA, dB, dC, extra):
    st = fn(0, 0, 0, M, N, K, ctypes.c_float(1.0), dA.ptr, K, dB.ptr, N,
            ctypes.c_float(0.0), dC.ptr, N, 0, *extra)
    return st


def _pair(hml, A, B, stale=None):
    """Runs hmlSgemmExNoAlloc and hmlSgemmEx (library workspace) on the same
    row-major inputs; returns both outputs."""
    M, K = A.shape
    N = B.shape[1]
    dA = hml.DeviceBuffer.from_numpy(A)
    dB = hml.DeviceBuffer.from_numpy(B)
    fill = np.full((M, N), np.nan, np.float32)
    dC1 = hml.DeviceBuffer.from_numpy(fill)
    dC2 = hml.DeviceBuffer.from_numpy(fill)
    print("no-alloc call %dx%dx%d" % (M, N, K), flush=True)
    st = _call(hml, hml.lib.hmlSgemmExNoAlloc, M, N, K, dA, dB, dC1, (None,))
    assert st == 0, hml.lib.hmlLastError()
    st = _call(hml, hml.lib.hmlSgemmEx, M, N, K, dA, dB, dC2, (None, 0, None))
    assert st == 0, hml.lib.hmlLastError()
    hml.synchronize()
    return dC1.to_numpy((M, N)), dC2.to_numpy((M, N))


def _need(hml, M, N, K):
    return hml.lib.hmlSgemmWorkspaceSize(0, 0, 0, M, N, K, None)


def test_no_alloc_matches_sgemm_ex_two_calls(hml):
    rng = np.random.default_rng(7)
    # the first call leaves partial sums of a larger problem in the scratch;
    # the second, smaller one must not see them
    shapes = [(8192, 16, 8192), (2048, 16, 2048)]
    for M, N, K in shapes:
        assert 0 < _need(hml, M, N, K) <= SCRATCH, (M, N, K)
    for M, N, K in shapes:
        A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
        B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
        c1, c2 = _pair(hml, A, B)
        assert np.isfinite(c1).all()
        assert np.array_equal(c1, c2), (M, N, K)
    for M, N, K in shapes:
        A = int_matrix(rng, (M, K))
        B = int_matrix(rng, (K, N))
        c1, c2 = _pair(hml, A, B)
        exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float32)
        assert np.array_equal(c1, exact), (M, N, K)
        assert np.array_equal(c2, exact), (M, N, K)


def test_no_alloc_gemv_route_and_dense(hml):
    rng = np.random.default_rng(8)
    # N == 1 goes to the SGEMV kernels (split rows: scratch partials); a
    # 2048^2 x 256 problem to the dense FP32 tiles (operand copies in the
    # scratch)
    for M, N, K in [(16384, 1, 16384), (2048, 2048, 256), (1000, 1, 3000)]:
        assert _need(hml, M, N, K) <= SCRATCH, (M, N, K)
        A = int_matrix(rng, (M, K))
        B = int_matrix(rng, (K, N))
        c1, c2 = _pair(hml, A, B)
        exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float32)
        assert np.array_equal(c1, exact), (M, N, K)
        assert np.array_equal(c2, exact), (M, N, K)


def test_no_alloc_rejects_oversized_scratch(hml):
    # 8192 x 8192 x 512 runs on the dense tiles, whose operand copies need
    # 32 MiB: more than the fixed scratch holds, so the call must refuse
    # before launching anything and leave C untouched
    M, N, K = 8192, 8192, 512
    assert _need(hml, M, N, K) > SCRATCH
    dA = hml.DeviceBuffer(4 * M * K)
    dB = hml.DeviceBuffer(4 * K * N)
    dA.fill_uniform(M * K, 1)
    dB.fill_uniform(K * N, 2)
    C0 = np.full((M, N), 3.0, np.float32)
    dC = hml.DeviceBuffer.from_numpy(C0)
    st = _call(hml, hml.lib.hmlSgemmExNoAlloc, M, N, K, dA, dB, dC, (None,))
    assert st == 7, st  # HML_STATUS_WORKSPACE_TOO_SMALL
    msg = hml.lib.hmlLastError().decode()
    assert "scratch" in msg and str(SCRATCH) in msg, msg
    hml.synchronize()
    assert np.array_equal(dC.to_numpy((M, N)), C0)
