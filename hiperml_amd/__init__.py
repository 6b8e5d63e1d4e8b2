"""hiperml_amd -- MI355X-native skinny SGEMM/SGEMV (HiperML hot path).

Python host mirror of the C ABI in include/.  The reference-compatible calls
``sgemm``/``sgemm_basic``/``sgemv``/``sgemv_basic`` take host (numpy) arrays
and behave like hmlSgemm/hmlSgemmBasic/hmlSgemv/hmlSgemvBasic
(include/hml_sgemm.h:21-41, include/hml_sgemv.h:19-37 of the reference):
row-major, C (or y) read even when beta == 0, result written back in place.
The ``*_ex`` calls take device buffers and return status codes as
exceptions.  Every call runs the HIP kernels of libhml.so; there is no CPU
fallback.
"""
import ctypes

import numpy as np

from ._lib import (GemmShape, SgemmVariant, SgemvVariant, TIMER_FN, lib,
                   EXPORTED, LIB_PATH)

ROW_MAJOR, COL_MAJOR = 0, 1
OP_N, OP_T = 0, 1
FLAG_READ_C_ALWAYS = 0x1

SGEMM_FAMILY = {"basic": 0, "dot": 1, "axpy": 2, "gemv": 3, "bf16x3": 4,
                "dense": 5}
SGEMV_FAMILY = {"basic": 0, "dot": 1, "axpy": 2}

STATUS = {0: "success", 1: "invalid value", 2: "HIP runtime error",
          3: "no kernel for this variant", 4: "allocation failed",
          5: "I/O error", 6: "unsupported arguments",
          7: "workspace too small", 8: "validation failed", 9: "parse error"}


class HmlError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s (%d): %s" % (STATUS.get(status, "?"), status, msg))
        self.status = status


def check(status):
    if status != 0:
        raise HmlError(status, lib.hmlLastError().decode(errors="replace"))


def _order(order):
    if order in (ROW_MAJOR, "row", "R"):
        return ROW_MAJOR
    if order in (COL_MAJOR, "col", "C"):
        return COL_MAJOR
    raise ValueError("order must be 'row' or 'col'")


def _op(t):
    return OP_T if t in (True, OP_T, "T", "t") else OP_N


class DeviceBuffer:
    """A device allocation owned by Python (hmlDeviceMalloc/hmlDeviceFree)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib.hmlDeviceMalloc(ctypes.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, arr):
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes)
        if arr.nbytes:
            check(lib.hmlMemcpyHtoD(buf.ptr, arr.ctypes.data, arr.nbytes))
        return buf

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        check(lib.hmlMemcpyHtoD(self.ptr, arr.ctypes.data, arr.nbytes))

    def to_numpy(self, shape, dtype=np.float32):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        if out.nbytes:
            check(lib.hmlMemcpyDtoH(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def fill_uniform(self, n, seed, lo=-1.0, hi=1.0, stream=None):
        assert 4 * n <= self.nbytes
        check(lib.hmlFillUniform(self.ptr, n, seed, lo, hi, stream))

    def free(self):
        if self.ptr:
            lib.hmlDeviceFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    check(lib.hmlDeviceSynchronize())


class Event:
    def __init__(self):
        p = ctypes.c_void_p()
        check(lib.hmlEventCreate(ctypes.byref(p)))
        self.ptr = p.value

    def record(self, stream=None):
        check(lib.hmlEventRecord(self.ptr, stream))

    def synchronize(self):
        check(lib.hmlEventSynchronize(self.ptr))

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        check(lib.hmlEventElapsedMs(ctypes.byref(ms), self.ptr, end.ptr))
        return ms.value

    def __del__(self):
        try:
            lib.hmlEventDestroy(self.ptr)
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = ctypes.c_void_p()
        check(lib.hmlStreamCreate(ctypes.byref(p)))
        self.ptr = p.value

    def synchronize(self):
        check(lib.hmlStreamSynchronize(self.ptr))

    def __del__(self):
        try:
            lib.hmlStreamDestroy(self.ptr)
        except Exception:
            pass


def flush_caches(stream=None):
    check(lib.hmlFlushCaches(stream))


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, DeviceBuffer):
        return x.ptr
    return int(x)


def _ld(order, rows, cols):
    """Dense leading dimension of a rows x cols matrix in `order`."""
    return max(1, cols if order == ROW_MAJOR else rows)


# ---- extended (device pointer) API ----------------------------------------

def sgemm_ex(M, N, K, A, B, C, alpha=1.0, beta=0.0, order="row",
             transA=False, transB=False, lda=None, ldb=None, ldc=None,
             flags=0, variant=None, workspace=None, ws_bytes=0, stream=None):
    """C = alpha*op(A)*op(B) + beta*C on device buffers (hmlSgemmEx)."""
    o, ta, tb = _order(order), _op(transA), _op(transB)
    if lda is None:
        lda = _ld(o, M, K) if ta == OP_N else _ld(o, K, M)
    if ldb is None:
        ldb = _ld(o, K, N) if tb == OP_N else _ld(o, N, K)
    if ldc is None:
        ldc = _ld(o, M, N)
    if variant is None:
        st = lib.hmlSgemmEx(o, ta, tb, M, N, K, alpha, _ptr(A), lda, _ptr(B),
                            ldb, beta, _ptr(C), ldc, flags, _ptr(workspace),
                            ws_bytes, stream)
    else:
        v = variant if isinstance(variant, SgemmVariant) else SgemmVariant(*variant)
        st = lib.hmlSgemmExVariant(o, ta, tb, M, N, K, alpha, _ptr(A), lda,
                                   _ptr(B), ldb, beta, _ptr(C), ldc, flags,
                                   ctypes.byref(v), _ptr(workspace), ws_bytes,
                                   stream)
    check(st)


def sgemv_ex(M, N, A, x, y, alpha=1.0, beta=0.0, order="row", trans=False,
             lda=None, flags=0, variant=None, workspace=None, ws_bytes=0,
             stream=None):
    """y = alpha*op(A)*x + beta*y on device buffers (hmlSgemvEx)."""
    o, t = _order(order), _op(trans)
    if lda is None:
        lda = _ld(o, M, N)
    if variant is None:
        st = lib.hmlSgemvEx(o, t, M, N, alpha, _ptr(A), lda, _ptr(x), 1, beta,
                            _ptr(y), 1, flags, _ptr(workspace), ws_bytes,
                            stream)
    else:
        v = variant if isinstance(variant, SgemvVariant) else SgemvVariant(*variant)
        st = lib.hmlSgemvExVariant(o, t, M, N, alpha, _ptr(A), lda, _ptr(x), 1,
                                   beta, _ptr(y), 1, flags, ctypes.byref(v),
                                   _ptr(workspace), ws_bytes, stream)
    check(st)


PRECISION = {"fp32": 0, "bf16x3": 1, "bf16x6": 2, "f16x3": 3}


KMEANS_SEPARATE_ROUNDING = 0x1
KMEANS_KERNEL = {None: 0, "screened": 0x2, "direct": 0x4}


def kmeans_assign(rows, ctrds, assign, num_rows, num_dims, num_clusts,
                  separate_rounding=False, kernel=None, stream=None):
    """assign[r] = nearest centroid of row r (hmlKmeansAssign) on device
    buffers: rows num_rows x num_dims, ctrds num_clusts x num_dims, both
    row-major float32; assign uint32.  kernel: None (by size), "screened"
    or "direct" (same indices)."""
    flags = (KMEANS_SEPARATE_ROUNDING if separate_rounding else 0) | KMEANS_KERNEL[kernel]
    check(lib.hmlKmeansAssign(_ptr(assign), _ptr(rows), num_rows, num_dims,
                              _ptr(ctrds), num_clusts, flags, stream))


def sgemm_dense(M, N, K, A, B, C, alpha=1.0, beta=0.0, precision="fp32",
                order="row", transA=False, transB=False, lda=None, ldb=None,
                ldc=None, flags=0, workspace=None, ws_bytes=0, stream=None):
    """Dense SGEMM on the matrix cores (hmlSgemmDense): precision 'fp32'
    (plain FP32 MFMA), 'bf16x3' or 'bf16x6' (split-precision bf16 MFMA)."""
    o, ta, tb = _order(order), _op(transA), _op(transB)
    if lda is None:
        lda = _ld(o, M, K) if ta == OP_N else _ld(o, K, M)
    if ldb is None:
        ldb = _ld(o, K, N) if tb == OP_N else _ld(o, N, K)
    if ldc is None:
        ldc = _ld(o, M, N)
    check(lib.hmlSgemmDense(PRECISION[precision], o, ta, tb, M, N, K, alpha,
                            _ptr(A), lda, _ptr(B), ldb, beta, _ptr(C), ldc,
                            flags, _ptr(workspace), ws_bytes, stream))


class GroupedPlan:
    """hmlSgemmGroupedCreate/Run/Destroy: many row-major problems
    C_i = alpha A_i B_i + beta C_i in one persistent launch.  `problems` is a
    list of (M, N, K, A, B, C) with device buffers (dense leading dims) or
    (M, N, K, A, lda, B, ldb, C, ldc)."""

    def __init__(self, problems, alpha=1.0, beta=0.0):
        n = len(problems)
        I64 = ctypes.c_int64 * n
        VP = ctypes.c_void_p * n
        Ms, Ns, Ks, As, las, Bs, lbs, Cs, lcs = ([] for _ in range(9))
        for pr in problems:
            if len(pr) == 6:
                M, N, K, A, B, C = pr
                la, lb, lc = K, N, N
            else:
                M, N, K, A, la, B, lb, C, lc = pr
            Ms.append(M); Ns.append(N); Ks.append(K)
            As.append(_ptr(A)); Bs.append(_ptr(B)); Cs.append(_ptr(C))
            las.append(la); lbs.append(lb); lcs.append(lc)
        h = ctypes.c_void_p()
        check(lib.hmlSgemmGroupedCreate(ctypes.byref(h), n, I64(*Ms), I64(*Ns),
                                        I64(*Ks), VP(*As), I64(*las), VP(*Bs),
                                        I64(*lbs), VP(*Cs), I64(*lcs), alpha,
                                        beta))
        self.ptr = h.value

    def run(self, graph=False, stream=None):
        check(lib.hmlSgemmGroupedRun(self.ptr, 1 if graph else 0, stream))

    def info(self):
        nt = ctypes.c_int64()
        nb = ctypes.c_int32()
        by = ctypes.c_size_t()
        check(lib.hmlSgemmGroupedInfo(self.ptr, ctypes.byref(nt), ctypes.byref(nb),
                                      ctypes.byref(by)))
        return {"tiles": nt.value, "blocks": nb.value, "device_bytes": by.value}

    KINDS = ("wave2", "wave1", "shared")
    CLASSES = 4  # by N: 1..16, 17..32, 33..48, larger

    @staticmethod
    def class_of(N):
        return (N + 15) // 16 - 1 if N <= 48 else 3

    def kinds(self):
        """Kernel kind name per tile class (None for an empty class)."""
        k = (ctypes.c_int32 * self.CLASSES)()
        check(lib.hmlSgemmGroupedKinds(self.ptr, k))
        return [self.KINDS[x] if x >= 0 else None for x in k]

    def set_kind(self, cls, kind):
        check(lib.hmlSgemmGroupedSetKind(self.ptr, cls, self.KINDS.index(kind)))

    def autotune(self, reps=5, stream=None):
        """hmlSgemmGroupedAutotune: times every kernel kind per class and
        keeps the fastest.  Returns {class: {kind: ms}} for the non-empty
        classes and the chosen kinds."""
        nk = len(self.KINDS)
        ms = (ctypes.c_float * (nk * self.CLASSES))()
        check(lib.hmlSgemmGroupedAutotune(self.ptr, reps, stream, ms))
        times = {c: {k: ms[nk * c + i] for i, k in enumerate(self.KINDS)}
                 for c in range(self.CLASSES) if ms[nk * c] >= 0}
        return {"ms": times, "kinds": self.kinds()}

    def destroy(self):
        if self.ptr:
            lib.hmlSgemmGroupedDestroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Mlp:
    """hmlMlp*: the learned selector's network, trained and evaluated by
    HIP kernels.  Arrays are numpy float32 (host)."""

    def __init__(self, n_in=6, hidden=64, n_hidden=3, n_out=None, seed=0,
                 _ptr=None):
        if _ptr is not None:
            self.ptr = _ptr
            return
        if n_out is None:
            n_out = lib.hmlSelectorClassCount()
        h = ctypes.c_void_p()
        check(lib.hmlMlpCreate(ctypes.byref(h), n_in, hidden, n_hidden, n_out, seed))
        self.ptr = h.value

    @classmethod
    def load(cls, path):
        h = ctypes.c_void_p()
        check(lib.hmlMlpLoad(ctypes.byref(h), path.encode()))
        return cls(_ptr=h.value)

    def save(self, path):
        check(lib.hmlMlpSave(self.ptr, path.encode()))

    def nparams(self):
        return lib.hmlMlpParamCount(self.ptr)

    def get_params(self):
        out = np.empty(self.nparams(), np.float32)
        check(lib.hmlMlpGetParams(self.ptr, _fp(out), out.size))
        return out

    def set_params(self, p):
        p = _f32(p)
        check(lib.hmlMlpSetParams(self.ptr, _fp(p), p.size))

    def loss_grad(self, X, T, mask):
        X, T, mask = _f32(X), _f32(T), _f32(mask)
        g = np.empty(self.nparams(), np.float32)
        loss = ctypes.c_float()
        check(lib.hmlMlpLossGrad(self.ptr, _fp(X), _fp(T), _fp(mask), X.shape[0],
                                 ctypes.byref(loss), _fp(g)))
        return loss.value, g

    def train(self, X, T, mask, epochs=10, batch=256, lr=1e-3, seed=0):
        X, T, mask = _f32(X), _f32(T), _f32(mask)
        losses = np.empty(epochs, np.float32)
        check(lib.hmlMlpTrain(self.ptr, _fp(X), _fp(T), _fp(mask), X.shape[0],
                              epochs, batch, lr, seed, _fp(losses)))
        return losses

    def predict(self, X, mask):
        X, mask = _f32(X), _f32(mask)
        P = np.empty(mask.shape, np.float32)
        check(lib.hmlMlpPredict(self.ptr, _fp(X), _fp(mask), X.shape[0], _fp(P)))
        return P

    def destroy(self):
        if self.ptr:
            lib.hmlMlpDestroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def _shape(M, N, K, order="row", transA=False, transB=False, beta0=True):
    return GemmShape(M, N, K, _order(order), _op(transA), _op(transB), 1 if beta0 else 0)


def selector_features(M, N, K, **kw):
    f = np.empty(lib.hmlSelectorFeatureCount(), np.float32)
    check(lib.hmlSelectorFeatures(ctypes.byref(_shape(M, N, K, **kw)), _fp(f)))
    return f


def selector_mask(M, N, K, **kw):
    m = np.empty(lib.hmlSelectorClassCount(), np.float32)
    check(lib.hmlSelectorMask(ctypes.byref(_shape(M, N, K, **kw)), _fp(m)))
    return m


def selector_variant(M, N, K, cls, **kw):
    v = SgemmVariant()
    st = lib.hmlSelectorVariant(ctypes.byref(_shape(M, N, K, **kw)), cls, ctypes.byref(v))
    return v if st == 0 else None


def selector_class_of(M, N, K, variant, **kw):
    v = variant if isinstance(variant, SgemmVariant) else SgemmVariant(*variant)
    return lib.hmlSelectorClassOf(ctypes.byref(_shape(M, N, K, **kw)), ctypes.byref(v))


def selector_install(net):
    check(lib.hmlSelectorInstall(net.ptr if net is not None else None))


def sgemm_select(M, N, K, order="row", transA=False, transB=False, beta=0.0):
    v = SgemmVariant()
    src = ctypes.c_int32()
    check(lib.hmlSgemmSelect(_order(order), _op(transA), _op(transB), M, N, K,
                             beta, ctypes.byref(v), ctypes.byref(src)))
    return v, src.value


def sgemv_select(M, N, order="row", trans=False, beta=0.0):
    v = SgemvVariant()
    src = ctypes.c_int32()
    check(lib.hmlSgemvSelect(_order(order), _op(trans), M, N, beta,
                             ctypes.byref(v), ctypes.byref(src)))
    return v, src.value


def check_f64(M, N, K, A, B, C, alpha=1.0, beta=0.0, C0=None, order="row",
              transA=False, transB=False, lda=None, ldb=None, ldc=None):
    """FP64 GPU check (hmlSgemmCheckF64).  Returns a dict of statistics:
    rel_fro = ||C-R||_F/||R||_F, bound_fro = ||alpha|A||B| + beta|C0|||_F /
    ||R||_F (multiply by gamma_K for the Higham bound), max_abs,
    max_rel_bound (max |C-R| / componentwise bound), nonfinite."""
    o, ta, tb = _order(order), _op(transA), _op(transB)
    if lda is None:
        lda = _ld(o, M, K) if ta == OP_N else _ld(o, K, M)
    if ldb is None:
        ldb = _ld(o, K, N) if tb == OP_N else _ld(o, N, K)
    if ldc is None:
        ldc = _ld(o, M, N)
    st = (ctypes.c_double * 6)()
    check(lib.hmlSgemmCheckF64(o, ta, tb, M, N, K, alpha, _ptr(A), lda,
                               _ptr(B), ldb, beta, _ptr(C0), _ptr(C), ldc, st))
    e2, r2, b2, emax, erel, nonfin = list(st)
    rn = r2 ** 0.5
    return {"rel_fro": (e2 ** 0.5) / rn if rn > 0 else float(e2 > 0),
            "bound_fro": (b2 ** 0.5) / rn if rn > 0 else 0.0,
            "err_fro": e2 ** 0.5, "ref_fro": rn, "abs_bound_fro": b2 ** 0.5,
            "max_abs": emax, "max_rel_bound": erel, "nonfinite": int(nonfin)}


def gamma(K, u=2.0 ** -24):
    """Higham's gamma_K = K u / (1 - K u)."""
    return K * u / (1.0 - K * u)


# ---- reference-compatible host-pointer API --------------------------------

def _host_f32(a, name):
    if not (isinstance(a, np.ndarray) and a.dtype == np.float32 and
            a.flags["C_CONTIGUOUS"]):
        raise TypeError("%s must be a C-contiguous float32 numpy array" % name)
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def sgemm(C, A, B, alpha=1.0, beta=0.0, basic=False):
    """hmlSgemm / hmlSgemmBasic: C (M x N) = alpha*A (M x K)*B (K x N) +
    beta*C, row-major host arrays, C updated in place."""
    M, K = A.shape
    K2, N = B.shape
    if K2 != K or C.shape != (M, N):
        raise ValueError("shape mismatch")
    fn = lib.hmlSgemmBasic if basic else lib.hmlSgemm
    fn(_host_f32(C, "C"), _host_f32(A, "A"), _host_f32(B, "B"), M, N, K,
       alpha, beta)
    return C


def sgemv(y, A, x, alpha=1.0, beta=0.0, basic=False):
    """hmlSgemv / hmlSgemvBasic: y (M) = alpha*A (M x N)*x (N) + beta*y."""
    M, N = A.shape
    if x.shape != (N,) or y.shape != (M,):
        raise ValueError("shape mismatch")
    fn = lib.hmlSgemvBasic if basic else lib.hmlSgemv
    fn(_host_f32(y, "y"), _host_f32(A, "A"), _host_f32(x, "x"), M, N, alpha,
       beta)
    return y


def config_load(path, kind="sgemm"):
    fn = lib.hmlSgemmConfigLoad if kind == "sgemm" else lib.hmlSgemvConfigLoad
    check(fn(path.encode()))


def config_save(path, kind="sgemm"):
    fn = lib.hmlSgemmConfigSave if kind == "sgemm" else lib.hmlSgemvConfigSave
    check(fn(path.encode()))


def config_clear():
    lib.hmlConfigClear()


def device_info(device=0):
    name = ctypes.create_string_buffer(64)
    cus = ctypes.c_int32()
    l2 = ctypes.c_int64()
    hbm = ctypes.c_int64()
    clk = ctypes.c_int32()
    check(lib.hmlDeviceInfo(device, name, ctypes.byref(cus), ctypes.byref(l2),
                            ctypes.byref(hbm), ctypes.byref(clk)))
    return {"arch": name.value.decode(), "cus": cus.value, "l2": l2.value,
            "hbm": hbm.value, "clock_khz": clk.value}


def device_count():
    return lib.hmlDeviceCount()


__all__ = [n for n in dir() if not n.startswith("_")]
