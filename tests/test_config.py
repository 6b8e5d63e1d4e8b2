"""Config store (host logic, no GPU): the v1.0 reader against known answers
from the reference's shipped tables, the v1.0 writer's exact format, the
v2.0 (MI355X) round trip, malformed input and the reference's fallback
rules (src/hml_sgemm_kernel.cu:189-237, src/hml_sgemv_kernel.cu:315-355).
"""
import ctypes
import json
import os

import pytest

import hiperml_amd as hml

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
REF_CFG = "/root/reference/config"
KAT = json.load(open(os.path.join(GOLDEN, "cfg_kat.json")))

SGEMM_HEADER_LINES = 18  # '#' lines of the v1.0 SGEMM header
SGEMV_HEADER_LINES = 18


def write_v1_from_kat(path, entry):
    with open(path, "w") as f:
        f.write("#SGEMM smart kernel config file v1.0\n" if entry["kind"] == "sgemm"
                else "#SGEMV smart kernel config file v1.0\n")
        for e in entry["entries"]:
            if entry["kind"] == "sgemm":
                f.write("%d %d : %d %d\n" % tuple(e))
            else:
                f.write("%d : %d %d %d\n" % tuple(e))


def cfg_source(name, tmp_path):
    """The shipped file when the reference is mounted, else a rebuild of
    its resolved table from the committed KAT."""
    p = os.path.join(REF_CFG, name)
    if os.path.exists(p):
        return p
    q = str(tmp_path / name)
    write_v1_from_kat(q, KAT[name])
    return q


def v1_get_sgemm(K, cs):
    t, r = ctypes.c_int(), ctypes.c_int()
    loaded = hml.lib.hmlSgemmConfigV1Get(K, cs, ctypes.byref(t), ctypes.byref(r))
    return loaded, t.value, r.value


def v1_get_sgemv(N):
    t, b, r = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    loaded = hml.lib.hmlSgemvConfigV1Get(N, ctypes.byref(t), ctypes.byref(b), ctypes.byref(r))
    return loaded, t.value, b.value, r.value


@pytest.fixture(autouse=True)
def clean():
    hml.config_clear()
    yield
    hml.config_clear()


@pytest.mark.parametrize("name", sorted(KAT))
def test_v1_reader_known_answers(name, tmp_path):
    entry = KAT[name]
    path = cfg_source(name, tmp_path)
    if entry["kind"] == "sgemm":
        hml.config_load(path, "sgemm")
        want = {(e[0], e[1]): (e[2], e[3]) for e in entry["entries"]}
        seen = 0
        for K in range(0, 97):
            for cs in range(0, 17):
                loaded, t, r = v1_get_sgemm(K, cs)
                assert loaded == 1
                exp = want.get((K, cs), (0, 0))
                assert (t, r) == exp, (K, cs)
                seen += (K, cs) in want
        assert seen == entry["unique_keys"] == 570
    else:
        hml.config_load(path, "sgemv")
        want = {e[0]: tuple(e[1:]) for e in entry["entries"]}
        for N in range(0, 1025):
            loaded, t, b, r = v1_get_sgemv(N)
            assert loaded == 1
            assert (t, b, r) == want.get(N, (0, 0, 0)), N
        assert entry["unique_keys"] == 1025


def test_kat_summary_counts():
    assert KAT["hml_sgemm_k40.cfg"]["type_histogram"] == {"0": 32, "1": 439, "2": 99}
    assert [0, 0, 10, 1] in KAT["hml_sgemv_k40.cfg"]["entries"]
    assert KAT["hml_sgemm_gtx480.cfg"]["rule_lines"] == 3 * 570


def test_v1_writer_format(tmp_path):
    # Loading the K40 table and printing it gives the reference header and
    # one "K colStops : type rowStops" line per valid rule in K-major order
    # (hmlSgemmKernelConfigPrint, src/hml_sgemm_kernel.cu:557-568).
    src = cfg_source("hml_sgemm_k40.cfg", tmp_path)
    hml.config_load(src, "sgemm")
    out = str(tmp_path / "out.cfg")
    hml.config_save(out, "sgemm")
    lines = open(out).read().splitlines()
    header = [l for l in lines if l.startswith("#")]
    assert header[0] == "#SGEMM smart kernel config file v1.0"
    assert len(header) == SGEMM_HEADER_LINES
    assert "#format: K colStops : kernelType rowStops" in header
    rules = [l for l in lines if not l.startswith("#")]
    want = ["%d %d : %d %d" % tuple(e) for e in KAT["hml_sgemm_k40.cfg"]["entries"] if e[3] > 0]
    assert rules == want
    # round trip
    hml.config_clear()
    hml.config_load(out, "sgemm")
    assert v1_get_sgemm(96, 1) == (1,) + tuple(
        next(e[2:] for e in KAT["hml_sgemm_k40.cfg"]["entries"] if e[:2] == [96, 1]))


def test_v1_sgemv_writer_format(tmp_path):
    src = cfg_source("hml_sgemv_k40.cfg", tmp_path)
    hml.config_load(src, "sgemv")
    out = str(tmp_path / "v.cfg")
    hml.config_save(out, "sgemv")
    lines = open(out).read().splitlines()
    assert lines[0] == "#SGEMV smart kernel config file v1.0"
    assert len([l for l in lines if l.startswith("#")]) == SGEMV_HEADER_LINES
    rules = [l for l in lines if not l.startswith("#")]
    assert rules[0] == "0 : 0 10 1"
    assert len(rules) == 1025


def test_malformed_and_out_of_range(tmp_path):
    p = tmp_path / "bad.cfg"
    p.write_text("#SGEMM smart kernel config file v1.0\n"
                 "2 1 : 2 8\n"
                 "\n"
                 "garbage line\n"
                 "500 1 : 1 4\n"        # K beyond the table: skipped
                 "3 40 : 1 4\n"         # colStops beyond the table: skipped
                 "4 2 : 1\n"            # too few fields: skipped
                 "2 1 : 1 5\n")         # duplicate: last one wins
    hml.config_load(str(p), "sgemm")
    assert v1_get_sgemm(2, 1) == (1, 1, 5)
    assert v1_get_sgemm(4, 2) == (1, 0, 0)
    with pytest.raises(hml.HmlError) as e:
        hml.config_load(str(tmp_path / "missing.cfg"), "sgemm")
    assert e.value.status == 5


def test_v1_selection_and_fallback(tmp_path):
    p = tmp_path / "r.cfg"
    p.write_text("16 1 : 1 8\n"    # valid var-K rule, rowStops 8 -> 4 row tiles
                 "32 1 : 0 6\n"    # basic
                 "64 1 : 2 0\n")   # invalid rule (rowStops 0) -> basic
    hml.config_load(str(p), "sgemm")
    v, src = hml.sgemm_select(100000, 16, 16)
    assert src == 1 and v.family == 1 and v.rt == 4 and v.ct == 1
    v, src = hml.sgemm_select(100000, 16, 32)
    assert src == 1 and (v.rt, v.ct, v.ksplit) == (1, 4, 1)
    v, src = hml.sgemm_select(100000, 16, 64)
    assert src == 1 and (v.rt, v.ct, v.ksplit) == (1, 4, 1)
    # K > 96 uses the K = 96 row (src/hml_sgemm_kernel.cu:206-211): absent -> basic
    v, src = hml.sgemm_select(100000, 16, 5000)
    assert src == 1 and (v.rt, v.ct, v.ksplit) == (1, 4, 1)
    # transposed layouts are not in the v1 tables: heuristic
    v, src = hml.sgemm_select(100000, 16, 16, transA=True)
    assert src == 0


def test_v1_sgemv_selection(tmp_path):
    p = tmp_path / "v.cfg"
    p.write_text("0 : 1 10 4\n64 : 2 8 2\n100 : 1 0 1\n")
    hml.config_load(str(p), "sgemv")
    v, src = hml.sgemv_select(100000, 64)
    assert src == 1 and v.family == 1 and v.rb == 2
    v, src = hml.sgemv_select(100000, 100)   # invalid blockStops -> basic
    assert src == 1 and (v.lpr, v.rb, v.split) == (64, 1, 1)
    v, src = hml.sgemv_select(1000, 5000)    # N > 1024 -> rule 0
    assert src == 1 and v.rb == 4


def test_v2_round_trip(tmp_path):
    s = hml.GemmShape(1 << 16, 16, 4096, 0, 0, 0, 1)
    v = hml.SgemmVariant(1, 2, 1, 4, 2)
    hml.check(hml.lib.hmlSgemmConfigAddRule(ctypes.byref(s), ctypes.byref(v)))
    s2 = hml.GemmShape(4096, 1, 4096, 0, 0, 0, 1)
    v2 = hml.SgemmVariant(3, 64, 4, 2, 4)
    hml.check(hml.lib.hmlSgemmConfigAddRule(ctypes.byref(s2), ctypes.byref(v2)))
    # replacing the same bucket keeps one rule
    v3 = hml.SgemmVariant(1, 1, 1, 8, 4)
    hml.check(hml.lib.hmlSgemmConfigAddRule(ctypes.byref(s), ctypes.byref(v3)))
    assert hml.lib.hmlSgemmConfigRuleCount() == 2
    out = str(tmp_path / "mi355x.cfg")
    hml.config_save(out, "sgemm")
    text = open(out).read()
    assert text.startswith("#SGEMM smart kernel config file v2.0")
    assert "0 16 4 12 : 1 1 1 8 4" in text
    hml.config_clear()
    hml.config_load(out, "sgemm")
    assert hml.lib.hmlSgemmConfigRuleCount() == 2
    got, src = hml.sgemm_select(1 << 16, 16, 4096)
    assert src == 1 and got.astuple() == (1, 1, 1, 8, 4)
    # nearest bucket of the same form
    got, src = hml.sgemm_select(1 << 17, 16, 8192)
    assert src == 1 and got.astuple() == (1, 1, 1, 8, 4)
    # a rule that cannot run its shape is rejected
    bad = hml.SgemmVariant(2, 1, 1, 1, 4)  # axpy family for a dot-form shape
    assert hml.lib.hmlSgemmConfigAddRule(ctypes.byref(s), ctypes.byref(bad)) == 3


def test_v2_sgemv_round_trip(tmp_path):
    v = hml.SgemvVariant(2, 16, 1, 256, 4)
    hml.check(hml.lib.hmlSgemvConfigAddRule(0, 1, 1 << 20, 64, ctypes.byref(v)))
    out = str(tmp_path / "v2.cfg")
    hml.config_save(out, "sgemv")
    hml.config_clear()
    hml.config_load(out, "sgemv")
    got, src = hml.sgemv_select(1 << 20, 64, trans=True)
    assert src == 1 and got.astuple() == (2, 16, 1, 256, 4)
    # the other family has no rule: heuristic
    got, src = hml.sgemv_select(1 << 20, 64)
    assert src == 0


def test_shipped_mi355x_configs_parse():
    root = os.path.join(os.path.dirname(__file__), "..", "config")
    for name, kind in (("hml_sgemm_mi355x.cfg", "sgemm"), ("hml_sgemv_mi355x.cfg", "sgemv")):
        p = os.path.join(root, name)
        if not os.path.exists(p):
            pytest.skip("MI355X config not generated yet")
        hml.config_load(p, kind)
        n = hml.lib.hmlSgemmConfigRuleCount() if kind == "sgemm" else hml.lib.hmlSgemvConfigRuleCount()
        assert n > 0
