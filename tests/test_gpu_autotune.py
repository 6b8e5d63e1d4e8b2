"""The exhaustive autotuner on the GPU, with its own timing (no fake timer):
every candidate is validated on the device and timed with hipEvents after a
cache flush (the counterpart of hmlSgemmKernelConfigOnline,
src/hml_sgemm_kernel.cu:356-555, which times without validating).  Checks:
every database row of the shape is a valid candidate with a positive
median, the returned winner is the database argmin, the heuristic's time is
the heuristic variant's row, a rule installed from the winner is what table
selection then returns, and a second run resumes from the database without
measuring anything."""
import csv
import ctypes
import time

import pytest

import hiperml_amd as hml

pytestmark = pytest.mark.gpu


def _rows(db, kind):
    out = []
    for r in csv.DictReader(open(db)):
        if int(r["kind"]) == kind:
            out.append(r)
    return out


def _var(r):
    return tuple(int(r[k]) for k in ("family", "p1", "p2", "p3", "p4"))


@pytest.mark.parametrize("shape", [(4096, 16, 2048, 0, 0), (3000, 5, 1000, 1, 0),
                                   (2048, 40, 512, 0, 1)],
                         ids=["4096x16x2048", "3000x5x1000-At", "2048x40x512-Bt"])
def test_sgemm_autotune_on_gpu(hml, tmp_path, shape):
    M, N, K, ta, tb = shape
    print("autotune", shape, flush=True)
    s = hml.GemmShape(M, N, K, 0, ta, tb, 1)
    db = str(tmp_path / "db.csv")
    best = hml.SgemmVariant()
    bt, ht = ctypes.c_double(), ctypes.c_double()
    st = hml.lib.hmlSgemmAutotune(ctypes.byref(s), 3, 1, 5, ctypes.byref(best),
                                  ctypes.byref(bt), ctypes.byref(ht), db.encode())
    assert st == 0, hml.lib.hmlLastError()
    arr = (hml.SgemmVariant * 2048)()
    n = hml.lib.hmlSgemmCandidates(ctypes.byref(s), arr, 2048)
    cands = {arr[i].astuple() for i in range(n)}
    rows = _rows(db, 0)
    assert {_var(r) for r in rows} == cands
    for r in rows:
        assert (int(r["M"]), int(r["N"]), int(r["K"])) == (M, N, K)
        # every candidate of these shapes runs and matches the reference
        assert r["valid"] == "1", r
        assert float(r["median_ms"]) > 0
    win = min(rows, key=lambda r: float(r["median_ms"]))
    assert _var(win) == best.astuple()
    assert bt.value == pytest.approx(float(win["median_ms"]), rel=1e-5)
    h, _ = hml.sgemm_select(M, N, K, transA=bool(ta), transB=bool(tb))
    hrow = [r for r in rows if _var(r) == h.astuple()]
    assert hrow and ht.value == pytest.approx(float(hrow[0]["median_ms"]), rel=1e-5)
    # the winner as a table rule: table selection returns it
    hml.config_clear()
    try:
        assert hml.lib.hmlSgemmConfigAddRule(ctypes.byref(s), ctypes.byref(best)) == 0
        v, src = hml.sgemm_select(M, N, K, transA=bool(ta), transB=bool(tb))
        assert src == 1 and v.astuple() == best.astuple()
    finally:
        hml.config_clear()
    # resume: nothing left to measure, so no rows are added and it is quick
    t0 = time.perf_counter()
    st = hml.lib.hmlSgemmAutotune(ctypes.byref(s), 3, 1, 5, ctypes.byref(best),
                                  ctypes.byref(bt), ctypes.byref(ht), db.encode())
    assert st == 0 and time.perf_counter() - t0 < 0.5
    assert len(_rows(db, 0)) == len(rows) and best.astuple() == _var(win)


def test_sgemv_autotune_on_gpu(hml, tmp_path):
    db = str(tmp_path / "v.csv")
    best = hml.SgemvVariant()
    bt, ht = ctypes.c_double(), ctypes.c_double()
    for trans in (0, 1):
        print("sgemv autotune trans", trans, flush=True)
        st = hml.lib.hmlSgemvAutotune(0, trans, 8192, 2048, 3, 1, 9, ctypes.byref(best),
                                      ctypes.byref(bt), ctypes.byref(ht), db.encode())
        assert st == 0, hml.lib.hmlLastError()
        rows = [r for r in _rows(db, 1) if int(r["transA"]) == trans]
        assert rows and all(r["valid"] == "1" and float(r["median_ms"]) > 0 for r in rows)
        win = min(rows, key=lambda r: float(r["median_ms"]))
        assert _var(win) == best.astuple()
        assert ht.value > 0 and bt.value <= ht.value
    out = str(tmp_path / "stats.txt")
    assert hml.lib.hmlAutotuneStats(db.encode(), 1, out.encode()) == 0
    assert "dot GFLOPS [order=0 trans=0]" in open(out).read()
