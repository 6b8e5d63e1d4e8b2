"""Host-side pieces of bench.py: the rocprofv3 counter parser and the
copy-kernel calibration of the HBM byte counters (no GPU)."""
import csv
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def write_pmc(path, counter, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Correlation_Id", "Dispatch_Id", "Grid_Size", "Kernel_Name",
                    "Counter_Name", "Counter_Value"])
        for i, (kn, g, kib) in enumerate(rows):
            w.writerow([i, i, g, kn, counter, kib])


def test_traffic_calibration(tmp_path):
    b = load_bench()
    calib_kib = b.CALIB_BYTES / 1024.0
    gemv = "void hml::k_gemv_dot_stream<64, 2>(hml::GemvParams)"
    copy = "__amd_rocclr_copyBuffer"
    fill = "hml::k_fill_uniform(float*, unsigned long, unsigned long, float, float)"
    # counters read 10% low for fetch, 5% high for write; the gemv kernel
    # moves 1.02x its algorithmic bytes
    alg = b.gemv_bytes(16384, 16384)
    fetch_true, write_true = 1.02 * alg, 4.0 * 16384
    write_pmc(tmp_path / "f.csv", "FETCH_SIZE",
              [(gemv, 1000, fetch_true * 0.9 / 1024)] * 3 +
              [(gemv, 10, 0.9 / 1024 * 4 * 2 ** 22)] +  # a smaller dispatch of the same kernel
              [(copy, 5, calib_kib * 0.9)] * 3 + [(fill, 7, 0.0)] * 3)
    write_pmc(tmp_path / "w.csv", "WRITE_SIZE",
              [(gemv, 1000, write_true * 1.05 / 1024)] * 3 +
              [(copy, 5, calib_kib * 1.05)] * 3 + [(fill, 7, calib_kib * 1.05)] * 3)
    out = b.traffic_from_pmc(str(tmp_path / "f.csv"), str(tmp_path / "w.csv"),
                             dst=str(tmp_path / "t.json"))
    assert out["fetch_factor"] == pytest.approx(1 / 0.9)
    assert out["write_factor"] == pytest.approx(1 / 1.05)
    assert out["traffic_bytes_per_launch"] == pytest.approx(fetch_true + write_true, rel=1e-9)
    assert out["kernel"] == gemv
    assert os.path.exists(tmp_path / "t.json")


def test_multi_gpu_request_exits():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "one GPU" in r.stderr
