"""The library workspace (hmlSgemmEx / hmlSgemvEx called without one) is kept
per (device, stream): calls queued on two streams at once never share
partial-sum buffers, and a buffer replaced by a larger one stays alive until
the calls using it are queued.  Integer-valued inputs make every result
exact, so any sharing or early free shows as a wrong bit."""
import numpy as np
import pytest

from util import int_matrix

pytestmark = pytest.mark.gpu


def _exact(A, B):
    return (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float32)


def test_two_streams_library_workspace(hml):
    rng = np.random.default_rng(21)
    # both shapes split K into partial sums (a workspace is needed)
    shapes = [(8192, 16, 8192), (4096, 32, 8192)]
    for M, N, K in shapes:
        assert hml.lib.hmlSgemmWorkspaceSize(0, 0, 0, M, N, K, None) > 0
    streams = [hml.Stream(), hml.Stream()]
    ops = []
    for (M, N, K) in shapes:
        A, B = int_matrix(rng, (M, K)), int_matrix(rng, (K, N))
        ops.append((M, N, K, hml.DeviceBuffer.from_numpy(A), hml.DeviceBuffer.from_numpy(B),
                    [hml.DeviceBuffer(4 * M * N) for _ in range(6)], _exact(A, B)))
    print("two-stream interleaved calls", flush=True)
    for rep in range(6):
        for s, (M, N, K, dA, dB, Cs, _) in zip(streams, ops):
            hml.sgemm_ex(M, N, K, dA, dB, Cs[rep], stream=s.ptr)
    for s in streams:
        s.synchronize()
    for M, N, K, dA, dB, Cs, want in ops:
        for C in Cs:
            np.testing.assert_array_equal(C.to_numpy((M, N)), want)


def test_growth_keeps_queued_buffers(hml):
    """Alternating small and large calls on one stream: every larger call
    replaces the stream's buffer while earlier calls may still be queued."""
    rng = np.random.default_rng(22)
    stream = hml.Stream()
    calls = []
    for M, N, K in [(1024, 8, 4096), (4096, 16, 8192), (2048, 4, 2048), (8192, 32, 8192),
                    (16384, 1, 16384)]:
        A, B = int_matrix(rng, (M, K)), int_matrix(rng, (K, N))
        dA, dB = hml.DeviceBuffer.from_numpy(A), hml.DeviceBuffer.from_numpy(B)
        dC = hml.DeviceBuffer(4 * M * N)
        hml.sgemm_ex(M, N, K, dA, dB, dC, stream=stream.ptr)
        calls.append((M, N, dC, _exact(A, B), dA, dB))
    stream.synchronize()
    for M, N, dC, want, *_ in calls:
        np.testing.assert_array_equal(dC.to_numpy((M, N)), want)


def test_large_stride_fallback_uses_caller_workspace(hml):
    """Row strides beyond the skinny kernels' 31-bit buffer offsets take the
    dense kernel, which copies its operands into the CALLER's workspace when
    one is given (or reports how much it needs)."""
    M, N, K = 16, 16, 64
    lda = 36_000_000  # 15 * lda floats > 2^31 bytes
    rng = np.random.default_rng(23)
    A = int_matrix(rng, (M, K))
    B = int_matrix(rng, (K, N))
    big = hml.DeviceBuffer(4 * ((M - 1) * lda + K))
    for i in range(M):
        hml.check(hml.lib.hmlMemcpyHtoD(big.ptr + 4 * i * lda, A[i].ctypes.data, 4 * K))
    dB = hml.DeviceBuffer.from_numpy(B)
    dC = hml.DeviceBuffer(4 * M * N)
    small = hml.DeviceBuffer(256)
    with pytest.raises(hml.HmlError) as e:
        hml.sgemm_ex(M, N, K, big, dB, dC, lda=lda, workspace=small, ws_bytes=256)
    assert e.value.status == 7 and "dense fallback" in str(e.value)
    ws = hml.DeviceBuffer(1 << 20)
    hml.sgemm_ex(M, N, K, big, dB, dC, lda=lda, workspace=ws, ws_bytes=1 << 20)
    hml.synchronize()
    np.testing.assert_array_equal(dC.to_numpy((M, N)), _exact(A, B))
