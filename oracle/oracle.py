"""Python wrapper of the CPU oracle (oracle/hml_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product (hiperml_amd) never imports it.

Functions and the reference code they restate:
  sgemm_cpu   apps/gemm/hml_sgemm_main.cu:260-292  (hmlSgemmCpu)
  sgemv_cpu   apps/gemv/hml_sgemv_main.cu:254-283  (hmlSgemvCpu)
  csum, ysum  apps/gemm/hml_sgemm_main.cu:294-305, apps/gemv/hml_sgemv_main.cu:285-292
  hml_gemm_inputs / hml_gemv_inputs  apps/gemm/hml_sgemm_main.cu:457-470,
              apps/gemv/hml_sgemv_main.cu:436-446, src/hml_utils.cu:60
  gemm_f64    FP64 oracle (double accumulation, any strided layout)
  flops_*     include/hml_flops.h:33-45,186,202
  kmeans_assign  include/hml_kmeans_kernel_template.h:18-88,90-316,323-367
              (GPU assign kernels, fused=True), src/hml_kmeans_cpu.cu:55-95
              (CPU path, fused=False)
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _load():
    if not os.path.exists(_SO):
        build()
    lib = ctypes.CDLL(_SO)
    fp = ctypes.POINTER(ctypes.c_float)
    dp = ctypes.POINTER(ctypes.c_double)
    i, f, l = ctypes.c_int, ctypes.c_float, ctypes.c_long
    lib.oracle_sgemm_cpu.argtypes = [fp, fp, fp, i, i, i, f, f]
    lib.oracle_sgemv_cpu.argtypes = [fp, fp, fp, i, i, f, f]
    lib.oracle_sgemm_cpu_rows.argtypes = [fp, fp, fp, i, i, i, f, f, i, i]
    lib.oracle_sgemv_cpu_rows.argtypes = [fp, fp, fp, i, i, f, f, i, i]
    lib.oracle_csum.argtypes = [fp, i, i]
    lib.oracle_csum.restype = ctypes.c_double
    lib.oracle_ysum.argtypes = [fp, i]
    lib.oracle_ysum.restype = ctypes.c_double
    lib.oracle_hml_random_fill.argtypes = [fp, ctypes.c_size_t, ctypes.c_uint, i]
    lib.oracle_hml_gemm_inputs.argtypes = [fp, fp, fp, i, i, i, ctypes.c_uint, i]
    lib.oracle_hml_gemv_inputs.argtypes = [fp, fp, fp, i, i, ctypes.c_uint, i]
    lib.oracle_gemm_f64.argtypes = [dp, dp, fp, fp, i, i, i, l, l, l, l]
    lib.oracle_flops_sgemm.argtypes = [ctypes.c_double] * 3
    lib.oracle_flops_sgemm.restype = ctypes.c_double
    lib.oracle_flops_sgemv.argtypes = [ctypes.c_double] * 2
    lib.oracle_flops_sgemv.restype = ctypes.c_double
    lib.oracle_kmeans_assign.argtypes = [ctypes.POINTER(ctypes.c_uint32), fp, l, i, fp, i, i]
    return lib


_lib = _load()
_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)


def _f(a):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_fp)


def sgemm_cpu(C, A, B, alpha=1.0, beta=0.0):
    """In place: C = alpha*A*B + beta*C, reference CPU order (row-major)."""
    M, K = A.shape
    _, N = B.shape
    _lib.oracle_sgemm_cpu(_f(C), _f(A), _f(B), M, N, K, alpha, beta)
    return C


def sgemm_cpu_rows(C, A, B, r0, r1, alpha=1.0, beta=0.0):
    M, K = A.shape
    _, N = B.shape
    _lib.oracle_sgemm_cpu_rows(_f(C), _f(A), _f(B), M, N, K, alpha, beta, r0, r1)
    return C


def sgemv_cpu(y, A, x, alpha=1.0, beta=0.0):
    M, N = A.shape
    _lib.oracle_sgemv_cpu(_f(y), _f(A), _f(x), M, N, alpha, beta)
    return y


def sgemv_cpu_rows(y, A, x, r0, r1, alpha=1.0, beta=0.0):
    M, N = A.shape
    _lib.oracle_sgemv_cpu_rows(_f(y), _f(A), _f(x), M, N, alpha, beta, r0, r1)
    return y


def csum(C):
    M, N = C.shape
    return _lib.oracle_csum(_f(C), M, N)


def ysum(y):
    return _lib.oracle_ysum(_f(y), y.shape[0])


def hml_random(n, seed=1, seed_given=False):
    out = np.empty(n, np.float32)
    _lib.oracle_hml_random_fill(_f(out), n, seed, 1 if seed_given else 0)
    return out


def hml_gemm_inputs(M, N, K, seed=1, seed_given=False):
    A = np.empty((M, K), np.float32)
    B = np.empty((K, N), np.float32)
    C = np.empty((M, N), np.float32)
    _lib.oracle_hml_gemm_inputs(_f(A), _f(B), _f(C), M, N, K, seed,
                                1 if seed_given else 0)
    return A, B, C


def hml_gemv_inputs(M, N, seed=1, seed_given=False):
    A = np.empty((M, N), np.float32)
    x = np.empty(N, np.float32)
    y = np.empty(M, np.float32)
    _lib.oracle_hml_gemv_inputs(_f(A), _f(x), _f(y), M, N, seed,
                                1 if seed_given else 0)
    return A, x, y


def gemm_f64(A, B, M, N, K, sAm, sAk, sBk, sBn, want_abs=False):
    """C64 (M x N, row-major float64) = sum_k A(m,k) B(k,n) in double, with
    A(m,k) = A.flat[m*sAm + k*sAk] and B(k,n) = B.flat[k*sBk + n*sBn]."""
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    C64 = np.empty((M, N), np.float64)
    absC = np.empty((M, N), np.float64) if want_abs else None
    _lib.oracle_gemm_f64(C64.ctypes.data_as(_dp),
                         absC.ctypes.data_as(_dp) if want_abs else None,
                         A.ctypes.data_as(_fp), B.ctypes.data_as(_fp),
                         M, N, K, sAm, sAk, sBk, sBn)
    return (C64, absC) if want_abs else C64


def flops_sgemm(M, N, K):
    return _lib.oracle_flops_sgemm(M, N, K)


def flops_sgemv(M, N):
    return _lib.oracle_flops_sgemv(M, N)


def layout_strides(order, trans, ld):
    """(stride of row index, stride of column index) of op(X) for a matrix
    stored in `order` ('row'/'col') with leading dimension ld."""
    rowview = (order == "row") == (not trans)
    return (ld, 1) if rowview else (1, ld)


def kmeans_assign(rows, ctrds, fused=True):
    """uint32 nearest-centroid index per row (numClusts where no distance is
    below FLT_MAX), in the reference kernels' arithmetic (fused=True) or the
    reference CPU path's (fused=False)."""
    rows = np.ascontiguousarray(rows, np.float32)
    ctrds = np.ascontiguousarray(ctrds, np.float32)
    n, d = rows.shape
    k, d2 = ctrds.shape
    assert d == d2
    out = np.empty(n, np.uint32)
    _lib.oracle_kmeans_assign(out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                              _f(rows), n, d, _f(ctrds), k, 1 if fused else 0)
    return out
