"""GPU parity tests of the skinny SGEMM path (hmlSgemm / hmlSgemmEx).

Oracles: the CPU restatement of the reference CPU path (oracle/hml_oracle.c,
apps/gemm/hml_sgemm_main.cu:260-292), the FP64 oracle (numpy float64 and the
GPU FP64 check kernel, itself checked against the CPU here), and closed
forms: integer-valued inputs make every partial sum exact, so the GPU result
must equal the exact product BIT FOR BIT whatever the summation order.
Floating-point tolerance: componentwise |C - R| <= 2 * gamma_{K+2} *
(|alpha||A||B| + |beta||C0|) (Higham, Accuracy and Stability, section 3.5).
"""
import ctypes
import itertools
import sys

import numpy as np
import pytest

from util import (assert_within_gamma, f64_ref, gamma, int_matrix, load_c,
                  store, store_c, uniform)

pytestmark = pytest.mark.gpu

LAYOUTS = [(o, ta, tb) for o in ("row", "col") for ta in (False, True)
           for tb in (False, True)]


def run_gemm(hml, A, B, C0, alpha, beta, order="row", ta=False, tb=False,
             variant=None, flags=0):
    M, K = A.shape
    N = B.shape[1]
    As, lda = store(A, order, ta)
    Bs, ldb = store(B, order, tb)
    Cs, ldc = store_c(C0, order)
    dA = hml.DeviceBuffer.from_numpy(As)
    dB = hml.DeviceBuffer.from_numpy(Bs)
    dC = hml.DeviceBuffer.from_numpy(Cs)
    hml.sgemm_ex(M, N, K, dA, dB, dC, alpha, beta, order=order, transA=ta,
                 transB=tb, lda=lda, ldb=ldb, ldc=ldc, variant=variant,
                 flags=flags)
    hml.synchronize()
    return load_c(dC, M, N, order)


def exact_product(A, B, alpha=1.0, beta=0.0, C0=None):
    R = A.astype(np.int64) @ B.astype(np.int64)
    R = alpha * R.astype(np.float64)
    if C0 is not None and beta != 0.0:
        R = R + beta * C0.astype(np.float64)
    return R.astype(np.float32)


SHAPES_INT = [(64, 16, 64), (1000, 7, 130), (33, 40, 17), (4096, 64, 1024),
              (16, 512, 300), (5, 3, 1000), (257, 1, 129), (1, 65, 33),
              (2048, 16, 2048), (300, 129, 70)]


@pytest.mark.parametrize("shape", SHAPES_INT, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("layout", LAYOUTS, ids=lambda l: "%s%s%s" % (l[0], "T" if l[1] else "N", "T" if l[2] else "N"))
def test_integer_exact_layouts(hml, shape, layout):
    M, N, K = shape
    order, ta, tb = layout
    print("case", shape, layout, flush=True)
    rng = np.random.default_rng(hash(shape) % 2**32)
    A = int_matrix(rng, (M, K))
    B = int_matrix(rng, (K, N))
    C0 = int_matrix(rng, (M, N))
    for alpha, beta in ((1.0, 0.0), (1.0, 1.0), (2.0, -3.0)):
        C = run_gemm(hml, A, B, C0, alpha, beta, order, ta, tb)
        np.testing.assert_array_equal(C, exact_product(A, B, alpha, beta, C0))


CAND_SHAPES = [(3000, 4, 700, "row", False, False),
               (2049, 8, 1000, "row", False, True),
               (4096, 16, 2048, "row", False, False),
               (8192, 64, 512, "row", False, False),
               (3000, 33, 700, "row", False, False),
               (4096, 16, 2048, "row", True, False),
               (2500, 40, 300, "col", False, False),
               (20000, 1, 512, "row", False, False),
               (20000, 1, 512, "row", True, False)]


@pytest.mark.parametrize("case", CAND_SHAPES, ids=lambda c: "%dx%dx%d-%s%s%s" % (c[0], c[1], c[2], c[3], "T" if c[4] else "N", "T" if c[5] else "N"))
def test_integer_exact_every_candidate(hml, case):
    """Every variant the autotuner may pick is bit-exact on integer data."""
    M, N, K, order, ta, tb = case
    shape = hml.GemmShape(M, N, K, 0 if order == "row" else 1, int(ta), int(tb), 0)
    cands = (hml.SgemmVariant * 512)()
    n = hml.lib.hmlSgemmCandidates(ctypes.byref(shape), cands, 512)
    assert 0 < n <= 512
    rng = np.random.default_rng(7)
    A = int_matrix(rng, (M, K))
    B = int_matrix(rng, (K, N))
    C0 = int_matrix(rng, (M, N))
    R = exact_product(A, B, 1.0, 1.0, C0)
    for i in range(n):
        v = cands[i]
        print("variant", v, flush=True)
        C = run_gemm(hml, A, B, C0, 1.0, 1.0, order, ta, tb, variant=v)
        np.testing.assert_array_equal(C, R, err_msg=repr(v))


DOTB_SHAPES = [(1000, 64, 700), (4097, 33, 1000), (300, 16, 64), (5000, 128, 2048),
               (64, 5, 63), (130, 48, 4100), (1, 1, 1), (777, 100, 129)]


@pytest.mark.parametrize("tb", [False, True], ids=["BN", "BT"])
def test_dotb_variants_exact(hml, tb):
    """The shared-B DOT kernel (family 7) for every instantiated tile shape,
    B N-contiguous (row-major K x N) and K-contiguous (transposed), one and
    three K chunks, ragged M / N / K: bit-exact on integer data."""
    rng = np.random.default_rng(17)
    for (M, N, K) in DOTB_SHAPES:
        A = int_matrix(rng, (M, K))
        B = int_matrix(rng, (K, N))
        C0 = int_matrix(rng, (M, N))
        R = exact_product(A, B, 1.0, 1.0, C0)
        for rt in (1, 2):
            for ct in (1, 2, 4):
                for ks in (1, 3):
                    v = hml.SgemmVariant(7, rt, ct, ks, 4)
                    print("case", M, N, K, "BT" if tb else "BN", v, flush=True)
                    C = run_gemm(hml, A, B, C0, 1.0, 1.0, "row", False, tb, variant=v)
                    np.testing.assert_array_equal(C, R, err_msg=repr(v))


AXPY_SHAPES = [(1000, 64, 700), (4097, 33, 1000), (301, 16, 64), (2050, 100, 1500),
               (64, 5, 63), (130, 48, 4100), (6, 3, 5), (4099, 2, 3000), (777, 100, 129)]
AXPY_TILES = [(1, 1), (1, 2), (1, 4), (1, 8), (2, 1), (2, 2), (2, 4)]


@pytest.mark.parametrize("tb", [False, True], ids=["BN", "BT"])
def test_axpy_variants_exact(hml, tb):
    """The AXPY-form kernel (A contiguous along M: row-major A^T) for every
    instantiated tile shape and workgroup size, K split inside one workgroup
    and across several, B N-contiguous (row-major K x N) and K-contiguous
    (transposed), M not a multiple of 4 (scalar A loads), K below one step
    and ragged: bit-exact on integer data."""
    rng = np.random.default_rng(23)
    for (M, N, K) in AXPY_SHAPES:
        A = int_matrix(rng, (M, K))
        B = int_matrix(rng, (K, N))
        C0 = int_matrix(rng, (M, N))
        R = exact_product(A, B, 1.0, 1.0, C0)
        As, lda = store(A, "row", True)
        Bs, ldb = store(B, "row", tb)
        dA = hml.DeviceBuffer.from_numpy(As)
        dB = hml.DeviceBuffer.from_numpy(Bs)
        for rt, ct in AXPY_TILES:
            for wpb in (1, 2, 4, 8):
                if wpb == 8 and rt * ct >= 8:
                    continue
                for ks in (1, 3, 8):
                    v = hml.SgemmVariant(2, rt, ct, ks, wpb)
                    print("case", M, N, K, "BT" if tb else "BN", v, flush=True)
                    dC = hml.DeviceBuffer.from_numpy(np.ascontiguousarray(C0))
                    hml.sgemm_ex(M, N, K, dA, dB, dC, 1.0, 1.0, order="row", transA=True,
                                 transB=tb, lda=lda, ldb=ldb, ldc=N, variant=v)
                    hml.synchronize()
                    np.testing.assert_array_equal(dC.to_numpy((M, N)), R, err_msg=repr(v))
                    dC.free()
        dA.free()
        dB.free()


GEMVN_SHAPES = [(1000, 2, 700, 0), (4097, 4, 1000, 3), (33, 8, 4100, 0), (300, 8, 63, 1),
                (64, 2, 5, 0), (2049, 4, 256, 0), (777, 2, 1030, 1)]


@pytest.mark.parametrize("tb", [False, True], ids=["BN", "BT"])
def test_gemvn_variants_exact(hml, tb):
    """The multi-vector GEMV family (8) for 2, 4 and 8 columns: every lanes-
    per-row / rows-per-group instantiation, no split, splits summed inside
    the workgroup and across workgroups, B N-contiguous and K-contiguous,
    ragged M and K, K below one step, padded lda (odd, breaking 16-byte
    alignment) and, with it, B rows not back to back (NaN in the gap):
    bit-exact on integer data."""
    rng = np.random.default_rng(29)
    for (M, N, K, pad) in GEMVN_SHAPES:
        A = int_matrix(rng, (M, K))
        B = int_matrix(rng, (K, N))
        C0 = int_matrix(rng, (M, N))
        R = exact_product(A, B, 1.0, 1.0, C0)
        lda = K + pad
        dA = hml.DeviceBuffer.from_numpy(padded(np.ascontiguousarray(A), lda))
        Bs, ldb = store(B, "row", tb)
        if pad % 2:  # B rows with a gap (not back to back)
            ldb += 1
            Bs = padded(Bs, ldb)
        dB = hml.DeviceBuffer.from_numpy(Bs)
        for lpr in (16, 32, 64):
            for rb in (1, 2, 4):
                for ks, wpb in ((1, 2), (3, 2), (8, 8), (16, 4)):
                    v = hml.SgemmVariant(8, lpr, rb, ks, wpb)
                    print("case", M, N, K, pad, ldb, "BT" if tb else "BN", v, flush=True)
                    dC = hml.DeviceBuffer.from_numpy(np.ascontiguousarray(C0))
                    hml.sgemm_ex(M, N, K, dA, dB, dC, 1.0, 1.0, order="row", transB=tb,
                                 lda=lda, ldb=ldb, ldc=N, variant=v)
                    hml.synchronize()
                    np.testing.assert_array_equal(dC.to_numpy((M, N)), R, err_msg=repr(v))
                    dC.free()
        dA.free()
        dB.free()


SMALL = [1, 2, 3, 5, 8, 13, 16, 17]
EDGE = [31, 32, 33, 63, 64, 65, 127, 128, 129, 255, 256, 257]


def small_shapes():
    out = list(itertools.product(SMALL, SMALL, SMALL))
    for e in EDGE:
        out += [(e, 7, 9), (9, e, 7), (7, 9, e), (e, e, 3), (e, 16, e)]
    return out


def test_bruteforce_small_vs_fp64(hml):
    """All M, N, K in a small set plus tile-edge sizes, random floats,
    against the FP64 oracle within the gamma_K bound."""
    rng = np.random.default_rng(11)
    for (M, N, K) in small_shapes():
        A = uniform(rng, (M, K))
        B = uniform(rng, (K, N))
        C0 = uniform(rng, (M, N))
        C = run_gemm(hml, A, B, C0, 1.5, 0.5)
        R, bound = f64_ref(A, B, 1.5, 0.5, C0)
        assert_within_gamma(C, R, bound, K)


def test_reference_cfg1_compat(hml):
    """Row-major 2048x2048x16 with the reference app's inputs
    (rand()/RAND_MAX, srand(1)), through the reference-compatible host API,
    against the restated reference CPU path and the FP64 oracle."""
    import oracle
    M, N, K = 2048, 16, 2048
    A, B, C = oracle.hml_gemm_inputs(M, N, K)
    Cgpu = C.copy()
    hml.sgemm(Cgpu, A, B, 1.0, 0.0)
    Ccpu = oracle.sgemm_cpu(C.copy(), A, B, 1.0, 0.0)
    R, bound = f64_ref(A, B)
    assert_within_gamma(Cgpu, R, bound, K)
    assert_within_gamma(Ccpu, R, bound, K)
    # The GPU sums in a different order; both stay inside the bound, and the
    # GPU is closer to FP64 than the sequential CPU loop.
    e_gpu = np.linalg.norm(Cgpu - R) / np.linalg.norm(R)
    e_cpu = np.linalg.norm(Ccpu - R) / np.linalg.norm(R)
    print("rel fro err gpu %.3e cpu %.3e" % (e_gpu, e_cpu))
    assert e_gpu <= e_cpu * 1.5 + 1e-12
    csum_gpu, csum_cpu = oracle.csum(Cgpu), oracle.csum(Ccpu)
    rel = abs(csum_gpu - csum_cpu) / csum_cpu
    print("Csum relative difference: %.20f" % rel)
    # apps/gemm/hml_sgemm_main.cu:33 gates at 1e-8; the two summation orders
    # differ by O(sqrt(K) u) per element, averaged over M*N entries.
    assert rel < 1e-6
    # Basic kernel too (hmlSgemmBasic).
    Cb = C.copy()
    hml.sgemm(Cb, A, B, 1.0, 0.0, basic=True)
    assert_within_gamma(Cb, R, bound, K)


@pytest.mark.parametrize("M,N,K", [(3000, 24, 5000), (70000, 300, 16)])
def test_compat_staged_copies(hml, monkeypatch, M, N, K):
    """hmlSgemm moves host operands of 4 MiB and more through pinned
    32 MiB chunks (A of 57 MiB: two chunks; C of 80 MiB: three, so both
    buffers are reused).  Integer-valued inputs: the result equals the
    exact product bit for bit, with staging on and with
    HML_HOST_STAGING=off (plain hipMemcpy)."""
    rng = np.random.default_rng(11)
    A = int_matrix(rng, (M, K))
    B = int_matrix(rng, (K, N))
    C0 = int_matrix(rng, (M, N))
    R = (A.astype(np.int64) @ B.astype(np.int64) * 2 - C0.astype(np.int64)).astype(np.float32)
    for mode in ("on", "off"):
        monkeypatch.setenv("HML_HOST_STAGING", mode)
        C = C0.copy()
        hml.sgemm(C, A, B, 2.0, -1.0)
        np.testing.assert_array_equal(C, R, err_msg="staging " + mode)


def test_compat_reads_c_when_beta_zero(hml):
    """The reference kernels always read C (beta*C with beta == 0), so a NaN
    in C propagates (include/hml_sgemm_kernel_template.h:187).  The
    compatible entry keeps that; the extended entry follows BLAS."""
    rng = np.random.default_rng(3)
    A = uniform(rng, (100, 50))
    B = uniform(rng, (50, 8))
    C0 = np.zeros((100, 8), np.float32)
    C0[3, 2] = np.nan
    C0[7, 5] = np.inf
    Cc = C0.copy()
    hml.sgemm(Cc, A, B, 1.0, 0.0)
    assert np.isnan(Cc[3, 2]) and np.isnan(Cc[7, 5])
    assert np.isfinite(np.delete(Cc.ravel(), [3 * 8 + 2, 7 * 8 + 5])).all()
    Cb = run_gemm(hml, A, B, C0, 1.0, 0.0)
    assert np.isfinite(Cb).all()
    Cs = run_gemm(hml, A, B, C0, 1.0, 0.0, flags=hml.FLAG_READ_C_ALWAYS)
    assert np.isnan(Cs[3, 2]) and np.isnan(Cs[7, 5])


def test_alpha_zero_and_k_zero(hml):
    rng = np.random.default_rng(5)
    A = uniform(rng, (64, 32))
    B = uniform(rng, (32, 16))
    C0 = uniform(rng, (64, 16))
    C = run_gemm(hml, A, B, C0, 0.0, 2.0)
    np.testing.assert_array_equal(C, (2.0 * C0.astype(np.float64)).astype(np.float32))
    C = run_gemm(hml, A, B, C0, 0.0, 0.0)
    np.testing.assert_array_equal(C, np.zeros_like(C0))
    A0 = np.zeros((64, 0), np.float32)
    B0 = np.zeros((0, 16), np.float32)
    dA = hml.DeviceBuffer(4)
    dB = hml.DeviceBuffer(4)
    dC = hml.DeviceBuffer.from_numpy(C0)
    hml.sgemm_ex(64, 16, 0, dA, dB, dC, 1.0, 0.5, lda=1, ldb=16, ldc=16)
    np.testing.assert_array_equal(dC.to_numpy((64, 16)),
                                  (0.5 * C0.astype(np.float64)).astype(np.float32))
    del A0, B0


def test_invalid_arguments(hml):
    with pytest.raises(hml.HmlError):
        hml.sgemm_ex(-1, 4, 4, None, None, None)
    d = hml.DeviceBuffer(1024)
    with pytest.raises(hml.HmlError):
        hml.sgemm_ex(16, 4, 8, d, d, d, lda=4)  # lda < K
    with pytest.raises(hml.HmlError):
        hml.sgemm_ex(16, 4, 8, d, d, d, variant=(1, 3, 1, 1, 4))  # no rt=3


def test_check_kernel_matches_cpu_fp64(hml):
    """The FP64 GPU check kernel agrees with the CPU FP64 oracle."""
    import oracle
    rng = np.random.default_rng(9)
    for order, ta, tb in LAYOUTS:
        M, N, K = 37, 11, 53
        A = uniform(rng, (M, K))
        B = uniform(rng, (K, N))
        C0 = uniform(rng, (M, N))
        # perturbed "result" with a known error
        R = (A.astype(np.float64) @ B.astype(np.float64)) * 0.75 + 0.25 * C0
        Cres = (R + 1e-3).astype(np.float32)
        As, lda = store(A, order, ta)
        Bs, ldb = store(B, order, tb)
        Cs, ldc = store_c(Cres, order)
        C0s, _ = store_c(C0, order)
        st = hml.check_f64(M, N, K, hml.DeviceBuffer.from_numpy(As),
                           hml.DeviceBuffer.from_numpy(Bs),
                           hml.DeviceBuffer.from_numpy(Cs), 0.75, 0.25,
                           C0=hml.DeviceBuffer.from_numpy(C0s), order=order,
                           transA=ta, transB=tb, lda=lda, ldb=ldb, ldc=ldc)
        sAm, sAk = oracle.layout_strides(order, ta, lda)
        sBk, sBn = oracle.layout_strides(order, tb, ldb)
        R64, absR = oracle.gemm_f64(As, Bs, M, N, K, sAm, sAk, sBk, sBn, want_abs=True)
        R64 = 0.75 * R64 + 0.25 * C0.astype(np.float64)
        err = Cres.astype(np.float64) - R64
        assert st["nonfinite"] == 0
        np.testing.assert_allclose(st["err_fro"], np.linalg.norm(err), rtol=1e-9)
        np.testing.assert_allclose(st["ref_fro"], np.linalg.norm(R64), rtol=1e-12)
        np.testing.assert_allclose(st["max_abs"], np.abs(err).max(), rtol=1e-9)
        bnd = 0.75 * absR + 0.25 * np.abs(C0.astype(np.float64))
        np.testing.assert_allclose(st["abs_bound_fro"], np.linalg.norm(bnd), rtol=1e-12)


# the benchmark sizes, the reference app's shape, and odd-K shapes whose rows
# miss 128-byte lines (the DOT kernel's cached-load path)
FULL = ([(8192, n, 8192) for n in (1, 2, 4, 8, 16, 32, 64, 128)] + [(2048, 16, 2048)] +
        [(8192, 16, 8191), (16384, 4, 3001), (65536, 32, 1000)])


@pytest.mark.parametrize("shape", FULL, ids=lambda s: "x".join(map(str, s)))
def test_full_size_vs_fp64_check(hml, shape):
    """Full benchmark sizes: every output element checked against FP64 on
    the GPU; the normwise error must sit inside gamma_K times the normwise
    bound (size-independent property)."""
    M, N, K = shape
    print("case", shape, flush=True)
    dA = hml.DeviceBuffer(4 * M * K)
    dB = hml.DeviceBuffer(4 * K * N)
    dC = hml.DeviceBuffer(4 * M * N)
    dA.fill_uniform(M * K, 101)
    dB.fill_uniform(K * N, 102)
    dC.fill_uniform(M * N, 103)
    hml.sgemm_ex(M, N, K, dA, dB, dC, 1.0, 0.0)
    st = hml.check_f64(M, N, K, dA, dB, dC)
    print("rel_fro %.3e bound*gamma %.3e max_rel_bound %.3e" %
          (st["rel_fro"], gamma(K) * st["bound_fro"], st["max_rel_bound"]))
    assert st["nonfinite"] == 0
    assert st["max_rel_bound"] <= 2 * gamma(K + 2)
    assert st["rel_fro"] <= gamma(K + 2) * st["bound_fro"]


def padded(S, ld):
    """Row-major storage S (rows x cols) copied into rows of length ld."""
    rows, cols = S.shape
    P = np.full((rows, ld), np.float32(np.nan))  # padding must never be read
    P[:, :cols] = S
    return P


@pytest.mark.parametrize("layout", LAYOUTS, ids=lambda l: "%s%s%s" % (l[0], "T" if l[1] else "N", "T" if l[2] else "N"))
def test_padded_leading_dims(hml, layout):
    """Leading dimensions larger than the stored extent (NaN in the padding,
    so any read of it shows up), odd paddings that break 16-byte alignment,
    integer data: bit-exact."""
    order, ta, tb = layout
    rng = np.random.default_rng(5)
    for (M, N, K, pad) in ((3000, 16, 700, 3), (777, 40, 2100, 37), (64, 5, 64, 1), (20000, 2, 300, 64)):
        print("case", M, N, K, pad, layout, flush=True)
        A = int_matrix(rng, (M, K))
        B = int_matrix(rng, (K, N))
        C0 = int_matrix(rng, (M, N))
        As, lda = store(A, order, ta)
        Bs, ldb = store(B, order, tb)
        Cs, ldc = store_c(C0, order)
        lda, ldb, ldc = lda + pad, ldb + pad, ldc + pad
        dA = hml.DeviceBuffer.from_numpy(padded(As, lda))
        dB = hml.DeviceBuffer.from_numpy(padded(Bs, ldb))
        Cp = padded(Cs, ldc)
        dC = hml.DeviceBuffer.from_numpy(Cp)
        hml.sgemm_ex(M, N, K, dA, dB, dC, 1.0, 1.0, order=order, transA=ta,
                     transB=tb, lda=lda, ldb=ldb, ldc=ldc)
        hml.synchronize()
        out = dC.to_numpy(Cp.shape)
        got = out[:, :Cs.shape[1]]
        if order == "col":
            got = got.T
        np.testing.assert_array_equal(got, exact_product(A, B, 1.0, 1.0, C0))
        # the padding of C is untouched
        assert np.isnan(out[:, Cs.shape[1]:]).all()


def test_large_stride_offsets(hml):
    """A row stride near the 31-bit byte-offset limit of the skinny kernel's
    buffer addressing (2 rows, lda = 2^27), and one past it for a full row
    tile (17 rows, lda = 2^26: the dense kernel takes over).  Integer data,
    bit-exact; only the valid rows are uploaded."""
    rng = np.random.default_rng(9)
    for (M, N, K, lda) in ((2, 2, 100, 1 << 27), (17, 16, 96, 1 << 26)):
        print("case", M, N, K, lda, flush=True)
        A = int_matrix(rng, (M, K))
        B = int_matrix(rng, (K, N))
        nfl = (M - 1) * lda + K
        dA = hml.DeviceBuffer(4 * nfl)
        hml.check(hml.lib.hmlMemsetD8(dA.ptr, 0, 4 * nfl))
        for i in range(M):
            row = np.ascontiguousarray(A[i])
            hml.check(hml.lib.hmlMemcpyHtoD(dA.ptr + 4 * i * lda, row.ctypes.data, row.nbytes))
        dB = hml.DeviceBuffer.from_numpy(B)
        dC = hml.DeviceBuffer(4 * M * N)
        hml.sgemm_ex(M, N, K, dA, dB, dC, lda=lda)
        hml.synchronize()
        np.testing.assert_array_equal(dC.to_numpy((M, N)), exact_product(A, B))
        dA.free()


def test_axpy_large_stride_offsets(hml):
    """AXPY form (row-major A^T, A(m,k) at k*lda + m) with a column stride
    just inside the per-step 31-bit offset limit of its buffer addressing
    (lda = 2^24) and one past it (lda = 2^25: the dense kernel takes over).
    Integer data, bit-exact; only the valid columns are uploaded."""
    rng = np.random.default_rng(10)
    for (M, N, K, lda) in ((70, 3, 20, 1 << 24), (70, 3, 20, 1 << 25)):
        print("case", M, N, K, lda, flush=True)
        A = int_matrix(rng, (M, K))
        B = int_matrix(rng, (K, N))
        nfl = (K - 1) * lda + M
        dA = hml.DeviceBuffer(4 * nfl)
        hml.check(hml.lib.hmlMemsetD8(dA.ptr, 0, 4 * nfl))
        for k in range(K):
            col = np.ascontiguousarray(A[:, k])
            hml.check(hml.lib.hmlMemcpyHtoD(dA.ptr + 4 * k * lda, col.ctypes.data, col.nbytes))
        dB = hml.DeviceBuffer.from_numpy(B)
        dC = hml.DeviceBuffer(4 * M * N)
        hml.sgemm_ex(M, N, K, dA, dB, dC, transA=True, lda=lda)
        hml.synchronize()
        np.testing.assert_array_equal(dC.to_numpy((M, N)), exact_product(A, B))
        dA.free()
        dB.free()
        dC.free()
