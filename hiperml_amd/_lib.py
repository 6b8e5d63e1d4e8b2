"""ctypes binding of libhml.so, the C-ABI of the MI355X skinny SGEMM/SGEMV
library (declarations in include/hml_sgemm.h, hml_sgemv.h, hml_ext.h).

The library is built in-tree by ``make -C hiperml_amd/csrc`` (or
``__graft_entry__.build()``).  There is no fallback: if the shared library is
missing, importing this module raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# HML_LIB names another build of the library (the host-sanitizer build of
# `make -C hiperml_amd/csrc check-asan`)
LIB_PATH = os.environ.get("HML_LIB") or os.path.join(_HERE, "libhml.so")

c_float_p = ctypes.POINTER(ctypes.c_float)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_void_p = ctypes.c_void_p
c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_u32 = ctypes.c_uint32
c_u64 = ctypes.c_uint64
c_size = ctypes.c_size_t


class SgemmVariant(ctypes.Structure):
    """hmlSgemmVariant (include/hml_ext.h)."""
    _fields_ = [("family", c_i32), ("rt", c_i32), ("ct", c_i32),
                ("ksplit", c_i32), ("wpb", c_i32)]

    def astuple(self):
        return (self.family, self.rt, self.ct, self.ksplit, self.wpb)

    def __repr__(self):
        return "SgemmVariant%r" % (self.astuple(),)


class SgemvVariant(ctypes.Structure):
    """hmlSgemvVariant (include/hml_ext.h)."""
    _fields_ = [("family", c_i32), ("lpr", c_i32), ("rb", c_i32),
                ("split", c_i32), ("wpb", c_i32)]

    def astuple(self):
        return (self.family, self.lpr, self.rb, self.split, self.wpb)

    def __repr__(self):
        return "SgemvVariant%r" % (self.astuple(),)


class GemmShape(ctypes.Structure):
    """hmlGemmShape (include/hml_ext.h)."""
    _fields_ = [("M", c_i64), ("N", c_i64), ("K", c_i64), ("order", c_i32),
                ("transA", c_i32), ("transB", c_i32), ("beta0", c_i32)]


TIMER_FN = ctypes.CFUNCTYPE(ctypes.c_double, c_i32, ctypes.POINTER(c_i64),
                            ctypes.POINTER(c_i32), c_void_p)

# name -> (restype, argtypes)
_SIGS = {
    # reference-compatible API (host pointers, exit on error)
    "hmlSgemmKernelConfig": (None, [ctypes.c_char_p]),
    "hmlSgemmBasic": (None, [c_float_p, c_float_p, c_float_p, ctypes.c_int,
                             ctypes.c_int, ctypes.c_int, ctypes.c_float,
                             ctypes.c_float]),
    "hmlSgemm": (None, [c_float_p, c_float_p, c_float_p, ctypes.c_int,
                        ctypes.c_int, ctypes.c_int, ctypes.c_float,
                        ctypes.c_float]),
    "hmlSgemvKernelConfig": (None, [ctypes.c_char_p]),
    "hmlSgemvBasic": (None, [c_float_p, c_float_p, c_float_p, ctypes.c_int,
                             ctypes.c_int, ctypes.c_float, ctypes.c_float]),
    "hmlSgemv": (None, [c_float_p, c_float_p, c_float_p, ctypes.c_int,
                        ctypes.c_int, ctypes.c_float, ctypes.c_float]),
    # extended API
    "hmlSgemmEx": (c_i32, [c_i32, c_i32, c_i32, c_i64, c_i64, c_i64,
                           ctypes.c_float, c_void_p, c_i64, c_void_p, c_i64,
                           ctypes.c_float, c_void_p, c_i64, c_u32, c_void_p,
                           c_size, c_void_p]),
    "hmlSgemmExNoAlloc": (c_i32, [c_i32, c_i32, c_i32, c_i64, c_i64, c_i64,
                                  ctypes.c_float, c_void_p, c_i64, c_void_p,
                                  c_i64, ctypes.c_float, c_void_p, c_i64,
                                  c_u32, c_void_p]),
    "hmlSgemmExVariant": (c_i32, [c_i32, c_i32, c_i32, c_i64, c_i64, c_i64,
                                  ctypes.c_float, c_void_p, c_i64, c_void_p,
                                  c_i64, ctypes.c_float, c_void_p, c_i64,
                                  c_u32, ctypes.POINTER(SgemmVariant),
                                  c_void_p, c_size, c_void_p]),
    "hmlSgemmWorkspaceSize": (c_size, [c_i32, c_i32, c_i32, c_i64, c_i64,
                                       c_i64, ctypes.POINTER(SgemmVariant)]),
    "hmlSgemmSelect": (c_i32, [c_i32, c_i32, c_i32, c_i64, c_i64, c_i64,
                               ctypes.c_float, ctypes.POINTER(SgemmVariant),
                               ctypes.POINTER(c_i32)]),
    "hmlSgemmDense": (c_i32, [c_i32, c_i32, c_i32, c_i32, c_i64, c_i64, c_i64,
                              ctypes.c_float, c_void_p, c_i64, c_void_p, c_i64,
                              ctypes.c_float, c_void_p, c_i64, c_u32, c_void_p,
                              c_size, c_void_p]),
    "hmlSgemmDenseWorkspaceSize": (c_size, [c_i32, c_i64, c_i64, c_i64]),
    "hmlSgemmGroupedCreate": (c_i32, [ctypes.POINTER(c_void_p), c_i32,
                                      ctypes.POINTER(c_i64), ctypes.POINTER(c_i64),
                                      ctypes.POINTER(c_i64), ctypes.POINTER(c_void_p),
                                      ctypes.POINTER(c_i64), ctypes.POINTER(c_void_p),
                                      ctypes.POINTER(c_i64), ctypes.POINTER(c_void_p),
                                      ctypes.POINTER(c_i64), ctypes.c_float,
                                      ctypes.c_float]),
    "hmlSgemmGroupedRun": (c_i32, [c_void_p, c_i32, c_void_p]),
    "hmlSgemmGroupedInfo": (c_i32, [c_void_p, ctypes.POINTER(c_i64),
                                    ctypes.POINTER(c_i32), ctypes.POINTER(c_size)]),
    "hmlSgemmGroupedKinds": (c_i32, [c_void_p, ctypes.POINTER(c_i32)]),
    "hmlSgemmGroupedSetKind": (c_i32, [c_void_p, c_i32, c_i32]),
    "hmlSgemmGroupedAutotune": (c_i32, [c_void_p, c_i32, c_void_p,
                                        ctypes.POINTER(ctypes.c_float)]),
    "hmlSgemmGroupedDestroy": (c_i32, [c_void_p]),
    "hmlSgemvEx": (c_i32, [c_i32, c_i32, c_i64, c_i64, ctypes.c_float,
                           c_void_p, c_i64, c_void_p, c_i64, ctypes.c_float,
                           c_void_p, c_i64, c_u32, c_void_p, c_size,
                           c_void_p]),
    "hmlSgemvExVariant": (c_i32, [c_i32, c_i32, c_i64, c_i64, ctypes.c_float,
                                  c_void_p, c_i64, c_void_p, c_i64,
                                  ctypes.c_float, c_void_p, c_i64, c_u32,
                                  ctypes.POINTER(SgemvVariant), c_void_p,
                                  c_size, c_void_p]),
    "hmlSgemvWorkspaceSize": (c_size, [c_i32, c_i32, c_i64, c_i64,
                                       ctypes.POINTER(SgemvVariant)]),
    "hmlSgemvSelect": (c_i32, [c_i32, c_i32, c_i64, c_i64, ctypes.c_float,
                               ctypes.POINTER(SgemvVariant),
                               ctypes.POINTER(c_i32)]),
    "hmlSgemmConfigLoad": (c_i32, [ctypes.c_char_p]),
    "hmlSgemmConfigSave": (c_i32, [ctypes.c_char_p]),
    "hmlSgemvConfigLoad": (c_i32, [ctypes.c_char_p]),
    "hmlSgemvConfigSave": (c_i32, [ctypes.c_char_p]),
    "hmlConfigClear": (None, []),
    "hmlSgemmConfigV1Get": (ctypes.c_int, [ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_int),
                                           ctypes.POINTER(ctypes.c_int)]),
    "hmlSgemvConfigV1Get": (ctypes.c_int, [ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_int),
                                           ctypes.POINTER(ctypes.c_int),
                                           ctypes.POINTER(ctypes.c_int)]),
    "hmlSgemmConfigRuleCount": (c_i64, []),
    "hmlSgemvConfigRuleCount": (c_i64, []),
    "hmlAutotuneSetTimer": (None, [TIMER_FN, c_void_p]),
    "hmlSgemmCandidates": (c_i32, [ctypes.POINTER(GemmShape),
                                   ctypes.POINTER(SgemmVariant), c_i32]),
    "hmlSgemvCandidates": (c_i32, [c_i32, c_i32, c_i64, c_i64,
                                   ctypes.POINTER(SgemvVariant), c_i32]),
    "hmlSgemmAutotune": (c_i32, [ctypes.POINTER(GemmShape), c_i32, c_i32,
                                 c_u64, ctypes.POINTER(SgemmVariant),
                                 c_double_p, c_double_p, ctypes.c_char_p]),
    "hmlSgemvAutotune": (c_i32, [c_i32, c_i32, c_i64, c_i64, c_i32, c_i32,
                                 c_u64, ctypes.POINTER(SgemvVariant),
                                 c_double_p, c_double_p, ctypes.c_char_p]),
    "hmlSgemmConfigAddRule": (c_i32, [ctypes.POINTER(GemmShape),
                                      ctypes.POINTER(SgemmVariant)]),
    "hmlSgemvConfigAddRule": (c_i32, [c_i32, c_i32, c_i64, c_i64,
                                      ctypes.POINTER(SgemvVariant)]),
    "hmlMlpCreate": (c_i32, [ctypes.POINTER(c_void_p), c_i32, c_i32, c_i32, c_i32,
                             c_u64]),
    "hmlMlpDestroy": (c_i32, [c_void_p]),
    "hmlMlpParamCount": (c_i64, [c_void_p]),
    "hmlMlpGetParams": (c_i32, [c_void_p, c_float_p, c_i64]),
    "hmlMlpSetParams": (c_i32, [c_void_p, c_float_p, c_i64]),
    "hmlMlpLossGrad": (c_i32, [c_void_p, c_float_p, c_float_p, c_float_p, c_i32,
                               c_float_p, c_float_p]),
    "hmlMlpTrain": (c_i32, [c_void_p, c_float_p, c_float_p, c_float_p, c_i64, c_i32,
                            c_i32, ctypes.c_float, c_u64, c_float_p]),
    "hmlMlpPredict": (c_i32, [c_void_p, c_float_p, c_float_p, c_i64, c_float_p]),
    "hmlMlpSave": (c_i32, [c_void_p, ctypes.c_char_p]),
    "hmlMlpLoad": (c_i32, [ctypes.POINTER(c_void_p), ctypes.c_char_p]),
    "hmlSelectorFeatureCount": (c_i32, []),
    "hmlSelectorClassCount": (c_i32, []),
    "hmlSelectorFeatures": (c_i32, [ctypes.POINTER(GemmShape), c_float_p]),
    "hmlSelectorMask": (c_i32, [ctypes.POINTER(GemmShape), c_float_p]),
    "hmlSelectorVariant": (c_i32, [ctypes.POINTER(GemmShape), c_i32,
                                   ctypes.POINTER(SgemmVariant)]),
    "hmlSelectorClassOf": (c_i32, [ctypes.POINTER(GemmShape),
                                   ctypes.POINTER(SgemmVariant)]),
    "hmlSelectorInstall": (c_i32, [c_void_p]),
    "hmlFillUniform": (c_i32, [c_void_p, c_size, c_u64, ctypes.c_float,
                               ctypes.c_float, c_void_p]),
    "hmlSgemmCheckF64": (c_i32, [c_i32, c_i32, c_i32, c_i64, c_i64, c_i64,
                                 ctypes.c_float, c_void_p, c_i64, c_void_p,
                                 c_i64, ctypes.c_float, c_void_p, c_void_p,
                                 c_i64, c_double_p]),
    "hmlKmeansAssign": (c_i32, [c_void_p, c_void_p, c_i64, c_i32, c_void_p,
                                c_i32, c_u32, c_void_p]),
    "hmlInit": (c_i32, []),
    "hmlDeviceCount": (ctypes.c_int, []),
    "hmlSetDevice": (c_i32, [ctypes.c_int]),
    "hmlDeviceInfo": (c_i32, [ctypes.c_int, ctypes.c_char_p,
                              ctypes.POINTER(c_i32), ctypes.POINTER(c_i64),
                              ctypes.POINTER(c_i64), ctypes.POINTER(c_i32)]),
    "hmlDeviceMalloc": (c_i32, [ctypes.POINTER(c_void_p), c_size]),
    "hmlDeviceFree": (c_i32, [c_void_p]),
    "hmlMemcpyHtoD": (c_i32, [c_void_p, c_void_p, c_size]),
    "hmlMemcpyDtoH": (c_i32, [c_void_p, c_void_p, c_size]),
    "hmlMemcpyDtoD": (c_i32, [c_void_p, c_void_p, c_size]),
    "hmlMemsetD8": (c_i32, [c_void_p, ctypes.c_int, c_size]),
    "hmlDeviceSynchronize": (c_i32, []),
    "hmlStreamCreate": (c_i32, [ctypes.POINTER(c_void_p)]),
    "hmlStreamDestroy": (c_i32, [c_void_p]),
    "hmlStreamSynchronize": (c_i32, [c_void_p]),
    "hmlEventCreate": (c_i32, [ctypes.POINTER(c_void_p)]),
    "hmlEventDestroy": (c_i32, [c_void_p]),
    "hmlEventRecord": (c_i32, [c_void_p, c_void_p]),
    "hmlEventSynchronize": (c_i32, [c_void_p]),
    "hmlEventElapsedMs": (c_i32, [ctypes.POINTER(ctypes.c_float), c_void_p,
                                  c_void_p]),
    "hmlFlushCaches": (c_i32, [c_void_p]),
    "hmlReadProbe": (c_i32, [c_void_p, c_size, c_void_p, c_void_p]),
    "hmlAutotuneStats": (c_i32, [ctypes.c_char_p, c_i32, ctypes.c_char_p]),
    "hmlStatusString": (ctypes.c_char_p, [c_i32]),
    "hmlLastError": (ctypes.c_char_p, []),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libhml.so not built at %s; run `make -C hiperml_amd/csrc` or "
            "__graft_entry__.build()" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()
EXPORTED = tuple(_SIGS)
