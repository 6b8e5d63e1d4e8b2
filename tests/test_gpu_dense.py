"""GPU tests of the dense SGEMM kernels: plain FP32 MFMA and split-precision
bf16x3 / bf16x6 / scaled f16x3 MFMA (hmlSgemmDense).

Exact pins: integers with at most 8 significant bits are exact in bf16, so
every variant must reproduce their product bit for bit; 12-bit integers
split into hi + lo (bf16x3) or hi + mid (bf16x6) exactly, so the expected
results of the two product sets (bf16x3 drops lo*lo) are computed exactly in
integer arithmetic from a numpy restatement of the round-to-nearest-even
split.  Random data: error against FP64 within Higham's gamma_K bound for
FP32, and within the split's truncation bound for bf16x3.  The scaled fp16
split is pinned exactly by a one-term product set computed in numpy with
the same roundings, and must match or beat the plain FP32 kernel's error
against FP64 at the benchmark sizes on both input sets.
"""
import numpy as np
import pytest

from util import assert_within_gamma, f64_ref, gamma, int_matrix, load_c, store, store_c, uniform, loguniform_signed

pytestmark = pytest.mark.gpu

LAYOUTS = [(o, ta, tb) for o in ("row", "col") for ta in (False, True) for tb in (False, True)]


def bf16_rne(x):
    """float32 -> nearest bf16 value (ties to even), as float32."""
    u = x.astype(np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def split(x, terms):
    out = []
    r = x.astype(np.float32)
    for _ in range(terms):
        h = bf16_rne(r)
        out.append(h)
        r = (r - h).astype(np.float32)
    return out


def run_dense(hml, A, B, C0, alpha, beta, precision, order="row", ta=False, tb=False):
    M, K = A.shape
    N = B.shape[1]
    As, lda = store(A, order, ta)
    Bs, ldb = store(B, order, tb)
    Cs, ldc = store_c(C0, order)
    dA, dB, dC = (hml.DeviceBuffer.from_numpy(x) for x in (As, Bs, Cs))
    hml.sgemm_dense(M, N, K, dA, dB, dC, alpha, beta, precision=precision,
                    order=order, transA=ta, transB=tb, lda=lda, ldb=ldb, ldc=ldc)
    hml.synchronize()
    return load_c(dC, M, N, order)


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "bf16x6", "f16x3"])
@pytest.mark.parametrize("layout", LAYOUTS, ids=lambda l: "%s%s%s" % (l[0], "T" if l[1] else "N", "T" if l[2] else "N"))
def test_small_integers_exact(hml, precision, layout):
    """Integer-valued operands: every product and partial sum is exact, so
    each precision must return the exact result.  The shapes run both block
    heights of the two-plane kernels (128 rows up to a padded K of 1024,
    256 rows above: K = 1500)."""
    rng = np.random.default_rng(1)
    for (M, N, K) in ((128, 128, 32), (200, 300, 77), (5, 130, 1000), (257, 33, 129),
                      (300, 160, 1500)):
        A = int_matrix(rng, (M, K), -8, 8)
        B = int_matrix(rng, (K, N), -8, 8)
        C0 = int_matrix(rng, (M, N), -8, 8)
        C = run_dense(hml, A, B, C0, 1.0, -1.0, precision, *layout)
        want = (A.astype(np.int64) @ B.astype(np.int64) - C0.astype(np.int64)).astype(np.float32)
        np.testing.assert_array_equal(C, want)


@pytest.mark.parametrize("precision,terms", [("bf16x3", 2), ("bf16x6", 3)])
def test_split_product_sets_exact(hml, precision, terms):
    """12-bit integers: hi has 8 significant bits and the remainder fits the
    next term, so the kept products are known exactly; bf16x3 must equal
    hi*hi + hi*lo + lo*hi (no lo*lo), bf16x6 the full product."""
    rng = np.random.default_rng(2)
    # K = 3 keeps every partial sum of |products| below 2^24 (exact in FP32)
    M, N, K = 256, 128, 3
    A = rng.integers(-2047, 2048, (M, K)).astype(np.float32)
    B = rng.integers(-2047, 2048, (K, N)).astype(np.float32)
    C0 = np.zeros((M, N), np.float32)
    C = run_dense(hml, A, B, C0, 1.0, 0.0, precision)
    sa, sb = split(A, terms), split(B, terms)
    for s in sa + sb:  # every term is an integer here
        assert np.array_equal(s, np.round(s))
    ia = [s.astype(np.int64) for s in sa]
    ib = [s.astype(np.int64) for s in sb]
    if terms == 2:
        want = ia[0] @ ib[0] + ia[0] @ ib[1] + ia[1] @ ib[0]
        assert not np.array_equal(want, A.astype(np.int64) @ B.astype(np.int64))
    else:
        want = A.astype(np.int64) @ B.astype(np.int64)
    assert np.abs(want).max() < 2 ** 24
    np.testing.assert_array_equal(C, want.astype(np.float32))


def f16_split_reference(A, B, alpha):
    """numpy restatement of the scaled fp16 split for K = 1: row m of A and
    column n of B scaled by 2^e with max magnitude in [2^14, 2^15), split
    into fp16 hi + lo (round to nearest even), then
    C = alpha * 2^-(e_m + e_n) * fl32(hi*hi + fl32(lo*hi + hi*lo))."""
    def exps(X, axis):
        mx = np.abs(X).max(axis=axis).astype(np.float64)
        e = np.zeros(mx.shape, np.int64)
        nz = mx > 0
        e[nz] = 14 - np.floor(np.log2(mx[nz])).astype(np.int64)
        return e
    ea, eb = exps(A, 1), exps(B, 0)
    As = np.ldexp(A.astype(np.float64), ea[:, None]).astype(np.float32)
    Bs = np.ldexp(B.astype(np.float64), eb[None, :]).astype(np.float32)
    ah = As.astype(np.float16).astype(np.float32)
    al = (As - ah).astype(np.float16).astype(np.float32)
    bh = Bs.astype(np.float16).astype(np.float32)
    bl = (Bs - bh).astype(np.float16).astype(np.float32)
    hh = (ah @ bh).astype(np.float32)            # K = 1: single exact products
    corr = ((al @ bh).astype(np.float32) + (ah @ bl).astype(np.float32)).astype(np.float32)
    acc = (hh + corr).astype(np.float32)
    out = np.ldexp(acc.astype(np.float64), -(ea[:, None] + eb[None, :]))
    return (np.float32(alpha) * out.astype(np.float32)).astype(np.float32)


def test_f16x3_product_set_exact(hml):
    """K = 1, random floats over a wide range of magnitudes per row and
    column: every product the kernel keeps is exact in FP32, so its output
    is fixed bit for bit by the scaling exponents, the fp16 hi / lo split,
    the dropped lo*lo term and the separate correction accumulator."""
    rng = np.random.default_rng(9)
    M, N, K = 300, 200, 1
    # magnitudes 1e-10 .. 1e10 per row / column: every scaled product and
    # every unscaled result stays in FP32's normal range
    A = (rng.uniform(-1, 1, (M, K)) * 10.0 ** rng.uniform(-10, 10, (M, 1))).astype(np.float32)
    B = (rng.uniform(-1, 1, (K, N)) * 10.0 ** rng.uniform(-10, 10, (1, N))).astype(np.float32)
    A[3, 0] = 0.0
    B[0, 5] = 0.0
    C = run_dense(hml, A, B, np.zeros((M, N), np.float32), 1.0, 0.0, "f16x3")
    want = f16_split_reference(A, B, 1.0)
    np.testing.assert_array_equal(C, want)
    # lo*lo is really dropped: the exact product differs somewhere
    assert not np.array_equal(want, (A.astype(np.float64) @ B.astype(np.float64)).astype(np.float32))


def test_random_error_bounds(hml):
    rng = np.random.default_rng(3)
    M, N, K = 384, 256, 1000
    A = uniform(rng, (M, K))
    B = uniform(rng, (K, N))
    C0 = uniform(rng, (M, N))
    R, bound = f64_ref(A, B, 1.0, 0.5, C0)
    errs = {}
    for prec in ("fp32", "bf16x3", "bf16x6", "f16x3"):
        C = run_dense(hml, A, B, C0, 1.0, 0.5, prec)
        errs[prec] = np.linalg.norm(C - R) / np.linalg.norm(R)
        if prec == "fp32":
            assert_within_gamma(C, R, bound, K)
        elif prec == "bf16x6":
            # dropped terms are O(2^-24) per product: FP32-level bound
            assert_within_gamma(C, R, bound, K, slack=4.0)
        elif prec == "f16x3":
            # per operand |x - hi - lo| <= 2^-22 |x|, and |lo*lo| <= 2^-22 |ab|
            err = np.abs(C.astype(np.float64) - R)
            assert (err <= (3 * 2.0 ** -22 + 2 * gamma(K + 2)) * bound + 1e-30).all()
        else:
            # each product loses at most ~3 * 2^-16 of |a||b|
            err = np.abs(C.astype(np.float64) - R)
            assert (err <= (3 * 2.0 ** -16 + 2 * gamma(K + 2)) * bound + 1e-30).all()
    print("rel fro err", errs)
    assert errs["bf16x6"] < 4 * errs["fp32"] + 1e-9
    assert errs["bf16x3"] > errs["bf16x6"]
    assert errs["f16x3"] <= errs["fp32"]


@pytest.mark.parametrize("shape", [(8192, 8192, 256), (2048, 2048, 2048), (8192, 8192, 8192)],
                         ids=["8192x8192x256", "2048cube", "8192cube"])
@pytest.mark.parametrize("dist", ["uniform", "loguniform"])
def test_large_vs_fp64_check(hml, shape, dist):
    M, N, K = shape
    print("case", shape, dist, flush=True)
    rng = np.random.default_rng(4)
    if dist == "uniform":
        A = uniform(rng, (M, K)); B = uniform(rng, (K, N))
    else:
        A = loguniform_signed(rng, (M, K)); B = loguniform_signed(rng, (K, N))
    dA, dB = hml.DeviceBuffer.from_numpy(A), hml.DeviceBuffer.from_numpy(B)
    dC = hml.DeviceBuffer(4 * M * N)
    res = {}
    for prec in ("fp32", "bf16x6", "bf16x3", "f16x3"):
        hml.sgemm_dense(M, N, K, dA, dB, dC, precision=prec)
        st = hml.check_f64(M, N, K, dA, dB, dC)
        res[prec] = st
        assert st["nonfinite"] == 0
    print({k: (v["rel_fro"], v["max_rel_bound"]) for k, v in res.items()})
    assert res["fp32"]["max_rel_bound"] <= 2 * gamma(K + 2)
    assert res["bf16x6"]["max_rel_bound"] <= 4 * gamma(K + 2)
    assert res["bf16x3"]["max_rel_bound"] <= 3 * 2.0 ** -16 + 2 * gamma(K + 2)
    assert res["f16x3"]["max_rel_bound"] <= 3 * 2.0 ** -22 + 2 * gamma(K + 2)
    # the matched-error bar: relative Frobenius error against FP64 no worse
    # than the plain FP32 kernel's on the same inputs
    assert res["f16x3"]["rel_fro"] <= res["fp32"]["rel_fro"], res


def test_ex_routes_wide_shapes_to_dense(hml):
    v, src = hml.sgemm_select(4096, 1024, 512)
    assert v.family == 5
    rng = np.random.default_rng(6)
    A = int_matrix(rng, (300, 200))
    B = int_matrix(rng, (200, 400))
    dA, dB = hml.DeviceBuffer.from_numpy(A), hml.DeviceBuffer.from_numpy(B)
    dC = hml.DeviceBuffer(4 * 300 * 400)
    hml.sgemm_ex(300, 400, 200, dA, dB, dC)
    np.testing.assert_array_equal(dC.to_numpy((300, 400)),
                                  (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float32))
