"""Search and selection logic of the exhaustive autotuner with a fake
timing source (host only): the argmin wins, failed candidates are skipped
and recorded (src/hml_sgemm_kernel.cu:463-472 invalidates a failing kernel),
all-failing is an error, and every timing lands in the CSV database."""
import csv
import ctypes

import pytest

import hiperml_amd as hml


class FakeTimer:
    def __init__(self, fn):
        self.calls = []
        self.fn = fn
        self.cb = hml.TIMER_FN(self._call)

    def _call(self, kind, dims, var, user):
        d = (dims[0], dims[1], dims[2])
        v = tuple(var[i] for i in range(5))
        self.calls.append((kind, d, v))
        return self.fn(kind, d, v)

    def __enter__(self):
        hml.lib.hmlAutotuneSetTimer(self.cb, None)
        return self

    def __exit__(self, *a):
        hml.lib.hmlAutotuneSetTimer(hml.TIMER_FN(), None)


def sgemm_autotune(shape, db=None):
    best = hml.SgemmVariant()
    bt, ht = ctypes.c_double(), ctypes.c_double()
    st = hml.lib.hmlSgemmAutotune(ctypes.byref(shape), 3, 0, 1, ctypes.byref(best),
                                  ctypes.byref(bt), ctypes.byref(ht),
                                  db.encode() if db else None)
    return st, best, bt.value, ht.value


def candidates(shape):
    arr = (hml.SgemmVariant * 1024)()
    n = hml.lib.hmlSgemmCandidates(ctypes.byref(shape), arr, 1024)
    return [arr[i].astuple() for i in range(n)]


def test_candidates_cover_heuristic_and_are_runnable():
    for M, N, K, ta in ((8192, 16, 8192, 0), (1 << 20, 64, 64, 0), (4096, 1, 4096, 0),
                        (4096, 33, 777, 1), (100, 5000, 300, 0)):
        s = hml.GemmShape(M, N, K, 0, ta, 0, 1)
        c = candidates(s)
        assert len(c) == len(set(c)) > 1
        h, _ = hml.sgemm_select(M, N, K, transA=bool(ta))
        assert h.astuple() in c


def test_argmin_wins_and_db_records_all(tmp_path):
    s = hml.GemmShape(8192, 16, 8192, 0, 0, 0, 1)
    cands = candidates(s)
    # fake cost: prefers rt=2, ct=1, ksplit=4, wpb=8; one candidate "fails"
    def cost(kind, d, v):
        fam, rt, ct, ks, wpb = v
        if (rt, ks, wpb) == (4, 1, 4):
            return -1.0
        return 1.0 + abs(rt - 2) + abs(ct - 1) + abs(ks - 4) * 0.1 + (wpb != 8) * 0.01
    db = str(tmp_path / "db.csv")
    with FakeTimer(cost) as ft:
        st, best, bt, ht = sgemm_autotune(s, db)
    assert st == 0
    assert best.astuple() == (1, 2, 1, 4, 8)
    assert bt == pytest.approx(1.0)
    assert len(ft.calls) == len(cands)
    rows = list(csv.DictReader(open(db)))
    assert len(rows) == len(cands)
    failed = [r for r in rows if r["valid"] == "0"]
    assert all(float(r["median_ms"]) < 0 for r in failed)
    h, _ = hml.sgemm_select(8192, 16, 8192)
    assert ht == pytest.approx(cost(0, None, h.astuple()))


def test_all_candidates_fail():
    s = hml.GemmShape(4096, 8, 1024, 0, 0, 0, 1)
    with FakeTimer(lambda k, d, v: -1.0):
        st, best, bt, ht = sgemm_autotune(s)
    assert st == 8


def test_sgemv_autotune_fake(tmp_path):
    best = hml.SgemvVariant()
    bt, ht = ctypes.c_double(), ctypes.c_double()
    def cost(kind, d, v):
        assert kind == 1
        return 1.0 / v[3] + 0.001 * v[1]     # prefers larger split, small lpr
    with FakeTimer(cost):
        st = hml.lib.hmlSgemvAutotune(0, 1, 1 << 20, 64, 3, 0, 1, ctypes.byref(best),
                                      ctypes.byref(bt), ctypes.byref(ht), None)
    assert st == 0
    arr = (hml.SgemvVariant * 512)()
    n = hml.lib.hmlSgemvCandidates(0, 1, 1 << 20, 64, arr, 512)
    costs = [cost(1, None, arr[i].astuple()) for i in range(n)]
    assert bt.value == pytest.approx(min(costs))


def test_autotune_rejects_bad_shape():
    s = hml.GemmShape(0, 8, 8, 0, 0, 0, 1)
    with FakeTimer(lambda k, d, v: 1.0):
        st, *_ = sgemm_autotune(s)
    assert st == 1


def test_resume_skips_recorded_candidates(tmp_path):
    """The timing database is the checkpoint: a second run over the same
    shape measures nothing and returns the recorded winner; a run that was
    cut short measures only the missing candidates."""
    s = hml.GemmShape(8192, 16, 8192, 0, 0, 0, 1)
    cands = candidates(s)
    cost = lambda k, d, v: 1.0 + 0.01 * sum(v)
    db = str(tmp_path / "db.csv")
    with FakeTimer(cost) as ft:
        st, best, bt, ht = sgemm_autotune(s, db)
    assert st == 0 and len(ft.calls) == len(cands)
    with FakeTimer(lambda k, d, v: pytest.fail("re-measured %s" % (v,))) as ft:
        st2, best2, bt2, ht2 = sgemm_autotune(s, db)
    assert st2 == 0 and ft.calls == []
    assert best2.astuple() == best.astuple() and bt2 == pytest.approx(bt)
    assert ht2 == pytest.approx(ht)
    assert len(list(csv.DictReader(open(db)))) == len(cands)
    # an interrupted sweep: keep the header, the first 5 rows and half a row
    lines = open(db).read().splitlines(True)
    db2 = str(tmp_path / "db2.csv")
    open(db2, "w").write("".join(lines[:6]) + lines[6][:10])
    with FakeTimer(cost) as ft:
        st3, best3, bt3, _ = sgemm_autotune(s, db2)
    assert st3 == 0 and len(ft.calls) == len(cands) - 5
    assert best3.astuple() == best.astuple()
    # a different shape in the same database is measured in full
    s2 = hml.GemmShape(4096, 16, 8192, 0, 0, 0, 1)
    with FakeTimer(cost) as ft:
        assert sgemm_autotune(s2, db)[0] == 0
    assert len(ft.calls) == len(candidates(s2))


def test_sgemv_resume(tmp_path):
    db = str(tmp_path / "v.csv")
    best = hml.SgemvVariant()
    bt, ht = ctypes.c_double(), ctypes.c_double()
    for expect_calls in (True, False):
        with FakeTimer(lambda k, d, v: 1.0 + v[1] * 0.01) as ft:
            st = hml.lib.hmlSgemvAutotune(0, 0, 4096, 4096, 3, 0, 1, ctypes.byref(best),
                                          ctypes.byref(bt), ctypes.byref(ht), db.encode())
        assert st == 0
        assert bool(ft.calls) == expect_calls


def test_stats_tables(tmp_path):
    """Summary tables of the timing database (the counterpart of the
    reference's GFLOPS and max / min tables): best GFLOP/s per family for
    every (K, N) and the max / min spread."""
    db = str(tmp_path / "db.csv")
    # cost in ms: dot variants 1 ms .. 2 ms (spread 200%), others slower
    def cost(kind, d, v):
        base = 1.0 if v[0] == 1 else 4.0
        return base * (1.0 + (v[3] > 1))
    with FakeTimer(cost):
        for N in (16, 32):
            for K in (4096, 8192):
                assert sgemm_autotune(hml.GemmShape(8192, N, K, 0, 0, 0, 1), db)[0] == 0
    out = str(tmp_path / "stats.txt")
    assert hml.lib.hmlAutotuneStats(db.encode(), 0, out.encode()) == 0
    text = open(out).read()
    assert "dot GFLOPS [M=8192 order=0 transA=0 transB=0 beta0=1]" in text
    assert "dot max / min" in text and "best GFLOPS (family)" in text
    lines = text.splitlines()
    i = lines.index([l for l in lines if l.startswith("dot GFLOPS")][0])
    assert lines[i + 1].split() == ["K", "N=16", "N=32"]
    row = lines[i + 3].split()
    assert row[0] == "4096"
    # best dot time 1 ms: 2 * 8192 * 16 * 4096 / 1e6 GFLOP/s
    assert float(row[1]) == pytest.approx(2 * 8192 * 16 * 4096 / 1e6, rel=1e-3)
    j = lines.index("dot max / min [M=8192 order=0 transA=0 transB=0 beta0=1]:")
    assert lines[j + 3].split()[1] == "200%"
    assert hml.lib.hmlAutotuneStats(b"/nonexistent/db.csv", 0, None) == 5
