"""Pins the CPU oracle (oracle/hml_oracle.c) before anything is checked
against it.  Each check is chosen so that a plausible restatement mistake
fails it: wrong loop order or accumulator type, FMA contraction, a
transposed operand, a dropped alpha/beta term, a wrong input stream.
"""
import math
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_glibc_rand_kat():
    # The first outputs of glibc rand() after srand(1) are a published
    # constant sequence; hmlRandomFloat divides by RAND_MAX = 2^31 - 1
    # (src/hml_utils.cu:60).
    first = [1804289383, 846930886, 1681692777, 1714636915, 1957747793]
    got = oracle.hml_random(5)
    want = np.array([np.float32(np.float32(v) / np.float32(2147483647)) for v in first])
    np.testing.assert_array_equal(got, want)
    # default state == srand(1); srand(2) differs
    np.testing.assert_array_equal(oracle.hml_random(5, 1, True), got)
    assert not np.array_equal(oracle.hml_random(5, 2, True), got)


def test_gemm_inputs_order():
    # A, then B, then C from one stream (apps/gemm/hml_sgemm_main.cu:457-470)
    M, N, K = 3, 2, 4
    A, B, C = oracle.hml_gemm_inputs(M, N, K, 7, True)
    s = oracle.hml_random(M * K + K * N + M * N, 7, True)
    np.testing.assert_array_equal(A.ravel(), s[:M * K])
    np.testing.assert_array_equal(B.ravel(), s[M * K:M * K + K * N])
    np.testing.assert_array_equal(C.ravel(), s[M * K + K * N:])
    A, x, y = oracle.hml_gemv_inputs(3, 5, 9, True)
    s = oracle.hml_random(15 + 5 + 3, 9, True)
    np.testing.assert_array_equal(np.concatenate([A.ravel(), x, y]), s)


def test_sgemm_cpu_closed_forms():
    rng = np.random.default_rng(0)
    M, N, K = 7, 5, 11
    # all ones: every entry is K exactly
    C = oracle.sgemm_cpu(np.zeros((M, N), np.float32), np.ones((M, K), np.float32),
                         np.ones((K, N), np.float32))
    np.testing.assert_array_equal(C, np.full((M, N), K, np.float32))
    # identity: C = B (catches a transposed B or A)
    B = rng.standard_normal((K, N)).astype(np.float32)
    C = oracle.sgemm_cpu(np.zeros((K, N), np.float32), np.eye(K, dtype=np.float32), B)
    np.testing.assert_array_equal(C, B)
    # permutation: rows of B permuted
    P = np.eye(K, dtype=np.float32)[rng.permutation(K)]
    C = oracle.sgemm_cpu(np.zeros((K, N), np.float32), P, B)
    np.testing.assert_array_equal(C, P @ B)
    # integer data: exact, with alpha and beta terms
    A = rng.integers(-3, 4, (M, K)).astype(np.float32)
    Bi = rng.integers(-3, 4, (K, N)).astype(np.float32)
    C0 = rng.integers(-3, 4, (M, N)).astype(np.float32)
    C = oracle.sgemm_cpu(C0.copy(), A, Bi, 2.0, -3.0)
    want = 2 * (A.astype(np.int64) @ Bi.astype(np.int64)) - 3 * C0.astype(np.int64)
    np.testing.assert_array_equal(C, want.astype(np.float32))


def test_sgemm_cpu_float_ascending_accumulation():
    # Row [1, 2^-24, ..., 2^-24] times ones: a float accumulator adding in
    # ascending k rounds every 1 + 2^-24 back to 1 (ties to even), so the
    # result is exactly 1.  Double accumulation, pairwise or reversed order
    # would give more.
    K = 64
    a = np.full((1, K), 2.0 ** -24, np.float32)
    a[0, 0] = 1.0
    C = oracle.sgemm_cpu(np.zeros((1, 1), np.float32), a, np.ones((K, 1), np.float32))
    assert C[0, 0] == 1.0
    # reversed: the small terms accumulate first, then 1 is added
    C = oracle.sgemm_cpu(np.zeros((1, 1), np.float32), a[:, ::-1].copy(),
                         np.ones((K, 1), np.float32))
    assert C[0, 0] == np.float32(1.0 + (K - 1) * 2.0 ** -24)


def test_sgemm_cpu_no_fma_contraction():
    # a*b = 1 + 2^-11 + 2^-24 rounds to 1 + 2^-11 in float; adding
    # -(1 + 2^-11) then gives 0.  A fused multiply-add would keep 2^-24.
    a = np.float32(1.0 + 2.0 ** -12)
    A = np.array([[-(1.0 + 2.0 ** -11), a]], np.float32)
    B = np.array([[1.0], [a]], np.float32)
    C = oracle.sgemm_cpu(np.zeros((1, 1), np.float32), A, B)
    assert C[0, 0] == 0.0
    y = oracle.sgemv_cpu(np.zeros(1, np.float32), A, B[:, 0].copy())
    assert y[0] == 0.0


def test_sgemm_cpu_epilogue_rounding():
    # C = alpha*sum + beta*C with separate roundings (no FMA): pick values
    # where fma(alpha, sum, beta*C) differs from round(alpha*sum)+beta*C.
    a = np.float32(1.0 + 2.0 ** -12)
    A = np.array([[a]], np.float32)
    B = np.array([[1.0]], np.float32)
    C0 = np.array([[-(1.0 + 2.0 ** -11)]], np.float32)
    C = oracle.sgemm_cpu(C0.copy(), A, B, alpha=float(a), beta=1.0)
    assert C[0, 0] == 0.0


def test_sgemv_cpu_matches_sgemm_n1():
    rng = np.random.default_rng(1)
    A = rng.standard_normal((33, 47)).astype(np.float32)
    x = rng.standard_normal(47).astype(np.float32)
    y0 = rng.standard_normal(33).astype(np.float32)
    y = oracle.sgemv_cpu(y0.copy(), A, x, 1.5, -0.5)
    C = oracle.sgemm_cpu(y0[:, None].copy(), A, x[:, None].copy(), 1.5, -0.5)
    np.testing.assert_array_equal(y, C[:, 0])


@pytest.mark.parametrize("order", ["row", "col"])
@pytest.mark.parametrize("ta", [False, True])
@pytest.mark.parametrize("tb", [False, True])
def test_f64_oracle_layouts_vs_numpy(order, ta, tb):
    rng = np.random.default_rng(3)
    M, N, K = 9, 6, 13
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    As = A if (order == "row") == (not ta) else A.T.copy()
    Bs = B if (order == "row") == (not tb) else B.T.copy()
    lda = As.shape[1]
    ldb = Bs.shape[1]
    sAm, sAk = oracle.layout_strides(order, ta, lda)
    sBk, sBn = oracle.layout_strides(order, tb, ldb)
    C64, absC = oracle.gemm_f64(As, Bs, M, N, K, sAm, sAk, sBk, sBn, want_abs=True)
    np.testing.assert_allclose(C64, A.astype(np.float64) @ B.astype(np.float64), rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(absC, np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64)), rtol=1e-13)


def test_fp32_within_gamma_of_f64():
    # brute force on small shapes: the float path stays inside Higham's bound
    rng = np.random.default_rng(4)
    for M, N, K in [(1, 1, 1), (3, 2, 17), (8, 8, 300), (5, 1, 1000)]:
        A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
        B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
        C = oracle.sgemm_cpu(np.zeros((M, N), np.float32), A, B)
        R = A.astype(np.float64) @ B.astype(np.float64)
        bnd = np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64))
        g = K * 2.0 ** -24 / (1 - K * 2.0 ** -24)
        assert (np.abs(C - R) <= g * bnd + 1e-30).all()


def test_checksums():
    rng = np.random.default_rng(5)
    C = rng.integers(-1000, 1000, (40, 30)).astype(np.float32)
    assert oracle.csum(C) == float(C.astype(np.int64).sum())
    C = rng.standard_normal((40, 30)).astype(np.float32)
    s = 0.0
    for v in C.ravel():  # the reference's row-major double accumulation
        s += float(v)
    assert oracle.csum(C) == s
    assert abs(oracle.csum(C) - math.fsum(map(float, C.ravel()))) < 1e-12
    y = C[:, 0].copy()
    assert oracle.ysum(y) == sum(float(v) for v in y)


def test_flops_formulas():
    # include/hml_flops.h: FLOPS_SGEMM = 2MNK, FLOPS_SGEMV = 2MN + 2M
    assert oracle.flops_sgemm(2, 3, 4) == 48.0
    assert oracle.flops_sgemv(3, 4) == 30.0


def test_golden_vectors():
    """Committed fixtures (tests/golden/make_golden.py) still reproduce."""
    g = np.load(os.path.join(GOLDEN, "oracle_vectors.npz"), allow_pickle=False)
    A, B, C = oracle.hml_gemm_inputs(int(g["gemm_M"]), int(g["gemm_N"]), int(g["gemm_K"]), 1, False)
    np.testing.assert_array_equal(A, g["gemm_A"])
    np.testing.assert_array_equal(B, g["gemm_B"])
    np.testing.assert_array_equal(C, g["gemm_C0"])
    Cr = oracle.sgemm_cpu(C.copy(), A, B, 1.0, 0.0)
    np.testing.assert_array_equal(Cr, g["gemm_C_cpu"])
    assert oracle.csum(Cr) == float(g["gemm_csum"])
    Av, x, y = oracle.hml_gemv_inputs(int(g["gemv_M"]), int(g["gemv_N"]), 1, False)
    yr = oracle.sgemv_cpu(y.copy(), Av, x, 1.0, 0.0)
    np.testing.assert_array_equal(yr, g["gemv_y_cpu"])
    np.testing.assert_array_equal(oracle.hml_random(8), g["rand_first8"])


# ---- k-means assignment ---------------------------------------------------

def fused_flip_case(seed=3):
    """2-D inputs on which the fused (GPU kernel) and the separately rounded
    (CPU path) distance sums differ, built so that the expected assignments
    follow from exact float64 arithmetic.  Row x = 0; centroid 0 at distance
    exactly 9 (delta (3, 0)); centroid 1 with delta (d0, d1) whose separately
    rounded sum U is 9 while the fused sum F is below 9.  Fused: 1 (F < 9);
    separate: 0 (9 is not below 9)."""
    rng = np.random.default_rng(seed)
    f32, f64 = np.float32, np.float64
    for _ in range(200000):
        d0 = f32(rng.uniform(1.5, 2.5))
        sq0 = f32(f64(d0) * f64(d0))                 # one rounding (exact product)
        d1 = f32(math.sqrt(9.0 - float(sq0)))
        sq1 = f32(f64(d1) * f64(d1))
        # both sums are exact in float64 (operands within a few binades)
        U = f32(f64(sq0) + f64(sq1))
        F = f32(f64(sq0) + f64(d1) * f64(d1))
        assert math.fsum([float(sq0), float(f64(d1) * f64(d1))]) == f64(sq0) + f64(d1) * f64(d1)
        if U == f32(9.0) and F < f32(9.0):
            rows = np.zeros((1, 2), np.float32)
            ctrds = np.array([[-3.0, 0.0], [-d0, -d1]], np.float32)
            return rows, ctrds
    raise AssertionError("no flip case found")


def test_kmeans_assign_fused_vs_separate():
    rows, ctrds = fused_flip_case()
    assert oracle.kmeans_assign(rows, ctrds, fused=True).tolist() == [1]
    assert oracle.kmeans_assign(rows, ctrds, fused=False).tolist() == [0]


def test_kmeans_assign_closed_forms():
    rng = np.random.default_rng(4)
    D, K = 13, 9
    ctrds = rng.standard_normal((K, D)).astype(np.float32)
    ctrds[5] = ctrds[2]                              # duplicate: 2 must win
    pick = rng.integers(0, K, 500)
    rows = ctrds[pick]
    want = np.where(pick == 5, 2, pick)
    for fused in (True, False):
        np.testing.assert_array_equal(oracle.kmeans_assign(rows, ctrds, fused), want)
    # one centroid: everything goes to it; NaN / overflowing rows: none
    one = oracle.kmeans_assign(rows, ctrds[:1])
    assert (one == 0).all()
    bad = np.array([[np.nan] * D, [3e38] * D, [0.0] * D], np.float32)
    assert oracle.kmeans_assign(bad, ctrds).tolist()[:2] == [K, K]


def test_kmeans_assign_vs_float64_argmin():
    rng = np.random.default_rng(6)
    for (n, D, K) in ((3000, 64, 40), (2000, 3, 7), (500, 17, 200)):
        ctrds = (rng.standard_normal((K, D)) * 4).astype(np.float32)
        rows = (ctrds[rng.integers(0, K, n)] + rng.standard_normal((n, D)).astype(np.float32))
        d = ((rows[:, None, :].astype(np.float64) - ctrds[None].astype(np.float64)) ** 2).sum(-1)
        srt = np.sort(d, 1)
        sep = srt[:, 1] - srt[:, 0] > 1e-4 * srt[:, 1]   # no rounding can flip these
        assert sep.mean() > 0.9
        got = oracle.kmeans_assign(rows, ctrds)
        np.testing.assert_array_equal(got[sep], np.argmin(d, 1)[sep])
