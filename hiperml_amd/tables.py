"""Regenerates the shipped MI355X config tables (config/hml_sgemm_mi355x.cfg,
config/hml_sgemv_mi355x.cfg) with the exhaustive autotuner, on one GPU.

The tables are the v2.0 counterpart of the reference's per-GPU .cfg files
(written by its online autotune, src/hml_sgemm_kernel.cu:356-602):
  1. the online sweeps behind hmlSgemmKernelConfig(NULL) /
     hmlSgemvKernelConfig(NULL) (a power-of-two shape grid), resumable
     through $HML_AUTOTUNE_DB;
  2. one rule per benchmark shape bucket, autotuned on that exact shape:
     row-major SGEMM 8192 x N x 8192 for N = 1 .. 128 and SGEMV A x /
     A^T x for 16384 x 16384 and 1,048,576 x 64.
A loaded table's rule for a problem's own shape bucket outranks the learned
selector; the nearest rule of another bucket ranks below it.

  python -m hiperml_amd.tables [--out config] [--db autotune_tables.csv]
"""
import argparse
import ctypes
import os

from . import GemmShape, SgemmVariant, SgemvVariant, check, config_clear, config_save, lib

BENCH_SGEMM = [(8192, n, 8192) for n in (1, 2, 4, 8, 16, 32, 64, 128)]
BENCH_SGEMV = [(16384, 16384, t) for t in (0, 1)] + [(1 << 20, 64, t) for t in (0, 1)]


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.dirname(
        os.path.abspath(__file__))), "config"))
    ap.add_argument("--db", default="autotune_tables.csv",
                    help="timing database (the checkpoint: recorded shapes are not re-timed)")
    ap.add_argument("--no-sweep", action="store_true", help="benchmark shapes only")
    a = ap.parse_args(argv)
    os.environ["HML_SELECTOR"] = "off"
    os.environ["HML_AUTOTUNE_DB"] = a.db
    check(lib.hmlInit())
    config_clear()
    if not a.no_sweep:
        lib.hmlSgemmKernelConfig(None)
        print("SGEMM online sweep: %d rules" % lib.hmlSgemmConfigRuleCount(), flush=True)
        lib.hmlSgemvKernelConfig(None)
        print("SGEMV online sweep: %d rules" % lib.hmlSgemvConfigRuleCount(), flush=True)
    for M, N, K in BENCH_SGEMM:
        s = GemmShape(M, N, K, 0, 0, 0, 1)
        best, bt, ht = SgemmVariant(), ctypes.c_double(), ctypes.c_double()
        check(lib.hmlSgemmAutotune(ctypes.byref(s), 5, 1, 17, ctypes.byref(best),
                                   ctypes.byref(bt), ctypes.byref(ht), a.db.encode()))
        check(lib.hmlSgemmConfigAddRule(ctypes.byref(s), ctypes.byref(best)))
        print("SGEMM %dx%dx%d: %s %.4f ms (heuristic %.4f ms)"
              % (M, N, K, best.astuple(), bt.value, ht.value), flush=True)
    for M, N, t in BENCH_SGEMV:
        best, bt, ht = SgemvVariant(), ctypes.c_double(), ctypes.c_double()
        check(lib.hmlSgemvAutotune(0, t, M, N, 5, 1, 19, ctypes.byref(best), ctypes.byref(bt),
                                   ctypes.byref(ht), a.db.encode()))
        check(lib.hmlSgemvConfigAddRule(0, t, M, N, ctypes.byref(best)))
        print("SGEMV %dx%d trans=%d: %s %.4f ms (heuristic %.4f ms)"
              % (M, N, t, best.astuple(), bt.value, ht.value), flush=True)
    os.makedirs(a.out, exist_ok=True)
    config_save(os.path.join(a.out, "hml_sgemm_mi355x.cfg"), "sgemm")
    config_save(os.path.join(a.out, "hml_sgemv_mi355x.cfg"), "sgemv")
    print("wrote the tables into %s" % a.out, flush=True)


if __name__ == "__main__":
    main()
