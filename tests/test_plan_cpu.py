"""Host-side launch planning of the skinny kernels, checked through
hmlSgemmWorkspaceSize (no GPU needed).

Both skinny forms cut K into chunks of a multiple of their K step (DOT: the
64-long pipeline stage, AXPY: 16); the wpb waves of a workgroup (AXPY: the
largest power of two up to wpb) take consecutive chunks and sum them in
LDS, so partial sums go to the workspace only when the chunks span kg > 1
workgroups: kg * M * N floats (rounded up to 256 bytes)."""
import ctypes
import math

import pytest

import hiperml_amd as hml


def ws(M, N, K, v, ta=False):
    v = hml.SgemmVariant(*v)
    return hml.lib.hmlSgemmWorkspaceSize(0, 1 if ta else 0, 0, M, N, K, ctypes.byref(v))


def dot_partials(M, N, K, ksplit, wpb):
    kchunk = math.ceil(math.ceil(K / ksplit) / 64) * 64
    nchunks = math.ceil(K / kchunk)
    kw = max(1, min(wpb, nchunks))
    kg = math.ceil(nchunks / kw)
    return 0 if kg == 1 else (kg * M * N * 4 + 255) // 256 * 256


@pytest.mark.parametrize("M,N,K,ks,wpb", [
    (8192, 16, 8192, 12, 4),   # 12 chunks of 704, 3 workgroups
    (8192, 16, 8192, 4, 8),    # 4 chunks in one workgroup: no partials
    (8192, 16, 8192, 8, 4),    # 8 chunks, 2 workgroups
    (2048, 48, 2048, 64, 8),   # 32 chunks of 64 (the stage), 4 workgroups
    (1000, 7, 130, 3, 2),      # 3 chunks of 64 -> 3 chunks, 2 workgroups
    (100, 16, 63, 5, 4),       # K below one stage: one chunk
])
def test_dot_workspace(M, N, K, ks, wpb):
    assert ws(M, N, K, (1, 1, 1, ks, wpb)) == dot_partials(M, N, K, ks, wpb)


def axpy_partials(M, N, K, ksplit, wpb):
    kchunk = math.ceil(math.ceil(K / ksplit) / 16) * 16
    nchunks = math.ceil(K / kchunk)
    kw = max(1, min(wpb, nchunks))
    while kw & (kw - 1):
        kw -= 1
    kg = math.ceil(nchunks / kw)
    return 0 if kg == 1 else (kg * M * N * 4 + 255) // 256 * 256


@pytest.mark.parametrize("M,N,K,ks,wpb", [
    (4096, 16, 1000, 4, 4),    # 4 chunks of 256 in one workgroup: no partials
    (4096, 16, 1000, 16, 4),   # 16 chunks of 64, 4 workgroups
    (8192, 2, 8192, 12, 8),    # 12 chunks of 688, 2 workgroups of 8 waves
    (1000, 7, 130, 3, 2),      # 3 chunks of 48, 2 workgroups
    (1000, 7, 130, 3, 4),      # 3 chunks, workgroup of 2 waves (power of two)
    (100, 16, 15, 5, 4),       # K below one step: one chunk
])
def test_axpy_workspace(M, N, K, ks, wpb):
    # AXPY form (transposed A): B is read in place, so the workspace is the
    # partials of the workgroups along K only
    assert ws(M, N, K, (2, 1, 1, ks, wpb), ta=True) == axpy_partials(M, N, K, ks, wpb)


def test_dot_variant_limits():
    shape = hml.GemmShape(8192, 64, 4096, 0, 0, 0, 1)
    cands = (hml.SgemmVariant * 512)()
    n = hml.lib.hmlSgemmCandidates(ctypes.byref(shape), cands, 512)
    for i in range(n):
        v = cands[i]
        if v.family != 1:
            continue
        assert v.ct <= 4 and v.rt * v.ct <= 8 and v.wpb in (1, 2, 4, 8)


@pytest.mark.parametrize("M,N,K,ks", [(8192, 64, 8192, 8), (1000, 100, 129, 3), (300, 40, 64, 1)])
def test_dotb_workspace(M, N, K, ks):
    # shared-B DOT (family 7): every K chunk is a workgroup, so kg = chunks
    kchunk = math.ceil(math.ceil(K / ks) / 64) * 64
    kg = math.ceil(K / kchunk)
    want = 0 if kg == 1 else (kg * M * N * 4 + 255) // 256 * 256
    assert ws(M, N, K, (7, 2, 4, ks, 4)) == want


def test_heuristic_dense_gate():
    """The dense FP32 kernel (family 5, no K split) only when its 128 x 128
    tiles fill the device and K is small or N exceeds 512; otherwise the
    K-split skinny kernels (measured up to 2x faster there)."""
    sel = lambda M, N, K: hml.sgemm_select(M, N, K)[0].family
    assert sel(8192, 1024, 8192) == 5            # truly dense
    assert sel(16384, 256, 256) == 5             # small K, tiles fill
    assert sel(417, 512, 63616) != 5             # 16 tiles, long K
    assert sel(32768, 132, 46598) != 5           # N <= 512, long K
    assert sel(8192, 64, 8192) == 1              # skinny DOT
    assert sel(8192, 1, 8192) == 3               # GEMV


def test_gemv_heuristic_split_by_size():
    """GEMV rows are split only for A >= 512 MB (about 64K waves), and
    mid-size problems keep one pass (the split's reduction costs more)."""
    v = lambda M, N: hml.sgemm_select(M, 1, N)[0]
    big, mid = v(16384, 16384), v(8192, 8192)
    assert big.family == 3 and big.ksplit > 1
    assert mid.family == 3 and mid.ksplit == 1


def test_heuristic_fp32_bound_split():
    """N > 48: the K split stays inside the workgroup when the row tiles
    fill a third of a resident round (no reduction pass), but a short, long-K
    problem still splits across workgroups."""
    sel = lambda M, N, K: hml.sgemm_select(M, N, K)[0]
    big = sel(8192, 64, 8192)
    assert big.family == 1 and big.ksplit <= big.wpb
    thin = sel(1024, 64, 65536)
    assert thin.family == 1 and thin.ksplit > thin.wpb


@pytest.mark.parametrize("variant,slices", [
    ((1, 64, 2, 8, 8), 0),      # one workgroup covers each row: no partials
    ((1, 64, 2, 8, 2), 4),      # 8 chunks, 2 per workgroup: 4 slices
    ((1, 64, 2, 3, 2), 2),      # 3 chunks rounded up to 4: 2 slices
    ((1, 64, 2, 1, 4), 0),      # no split
    ((2, 32, 1, 8, 4), 8),      # AXPY: every split is a slice
])
def test_gemv_workspace_in_workgroup_sum(variant, slices):
    """SGEMV DOT splits: the waves of a workgroup sum their chunks in LDS,
    so the partials hold split / wpb slices of the 16384 outputs."""
    v = hml.SgemvVariant(*variant)
    trans = 1 if variant[0] == 2 else 0
    got = hml.lib.hmlSgemvWorkspaceSize(0, trans, 16384, 16384, ctypes.byref(v))
    assert got == (0 if slices <= 1 else slices * 16384 * 4)


def test_gemv_axpy_small_split_target():
    """Row-major A^T x below 64 MB of A aims at about 1024 workgroups
    (latency-bound short row ranges), larger problems at 512."""
    small = hml.sgemv_select(4099, 2049, trans=True)[0]
    big = hml.sgemv_select(16384, 16384, trans=True)[0]
    ncb = lambda n, lpr: -(-n // (4 * lpr))
    assert small.family == 2 and ncb(2049, small.lpr) * small.split >= 1000
    assert big.family == 2 and ncb(16384, big.lpr) * big.split <= 600


def test_heuristic_two_columns_multi_vector():
    """Two columns with A contiguous along K go to the multi-vector GEMV
    family (8) with the SGEMV geometry of the same rows; 4 and 8 columns
    stay on the matrix-core tiles, and the AXPY form is unaffected."""
    sel = lambda M, N, K, **kw: hml.sgemm_select(M, N, K, **kw)[0]
    v = sel(8192, 2, 8192)
    g = hml.sgemv_select(8192, 8192)[0]
    assert (v.family, v.rt, v.ct, v.ksplit, v.wpb) == (8, g.lpr, min(g.rb, 4), g.split, g.wpb)
    assert sel(8192, 2, 8192, transB=True).family == 8
    assert sel(8192, 4, 8192).family == 1
    assert sel(8192, 8, 8192).family == 1
    assert sel(8192, 2, 8192, transA=True).family == 2


@pytest.mark.parametrize("M,N,K,ks,wpb,slices", [
    (8192, 2, 8192, 1, 2, 0),     # no split
    (8192, 4, 8192, 8, 2, 4),     # 8 chunks, pairs summed in LDS: 4 slices
    (8192, 8, 8192, 3, 2, 2),     # 3 chunks rounded up to 4: 2 slices
    (1000, 2, 700, 8, 8, 0),      # one 8-wave workgroup covers the split
])
def test_gemvn_workspace(M, N, K, ks, wpb, slices):
    """Multi-vector GEMV family (8): partial slices of M x N floats follow
    the SGEMV DOT split rules (lanes per row 64 here)."""
    got = ws(M, N, K, (8, 64, 2, ks, wpb))
    assert got == (0 if slices <= 1 else (slices * M * N * 4 + 255) // 256 * 256)


def test_gemvn_runnable_and_selector_class():
    """Family 8 runs only for 2, 4 or 8 columns with A contiguous along K;
    selector class 86 is that family with the SGEMV heuristic's geometry."""
    assert hml.selector_variant(8192, 2, 8192, 86).astuple()[0] == 8
    assert hml.selector_variant(8192, 16, 8192, 86) is None          # not 2/4/8 columns
    assert hml.selector_variant(8192, 2, 8192, 86, transA=True) is None  # AXPY form
    m = hml.selector_mask(8192, 4, 8192)
    assert m[86] == 1
    v = hml.selector_variant(8192, 4, 8192, 86)
    assert hml.selector_class_of(8192, 4, 8192, v) == 86
