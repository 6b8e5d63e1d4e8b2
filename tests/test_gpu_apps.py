"""The CLI drivers (apps/gemm, apps/gemv) on the GPU: reference flags, the
reference's stdout lines, and its -c cross-check (CPU vs GPU checksum,
gate 1e-8, apps/gemm/hml_sgemm_main.cu:497-511)."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def app(name):
    path = os.path.join(ROOT, "apps", name)
    if not os.path.exists(path):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "apps")])
    return path


def write_cfg(tmp_path, kind):
    # an empty v2.0 table: no online autotune, heuristic selection
    p = tmp_path / ("%s.cfg" % kind)
    p.write_text("#%s smart kernel config file v2.0 (MI355X gfx950)\n" % kind.upper())
    return str(p)


def test_sgemm_app(tmp_path):
    cfg = write_cfg(tmp_path, "sgemm")
    out = tmp_path / "out.cfg"
    r = subprocess.run([app("gemm/hml_sgemm"), "-M", "4096", "-N", "16", "-K", "512",
                        "-c", "-r", "3", "-t", "3", "-s", cfg, "-o", str(out)],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0
    assert re.search(r"^GPU Csum = \d+\.\d{20}$", r.stdout, re.M)
    assert re.search(r"^CPU Csum = \d+\.\d{20}$", r.stdout, re.M)
    rel = float(re.search(r"Csum relative difference: ([0-9.e-]+)", r.stdout).group(1))
    assert rel <= 1e-8
    assert "Smart kernel median GFLOPS" in r.stderr
    assert out.read_text().startswith("#SGEMM smart kernel config file v2.0")


def test_sgemv_app(tmp_path):
    cfg = write_cfg(tmp_path, "sgemv")
    r = subprocess.run([app("gemv/hml_sgemv"), "-M", "20000", "-N", "300", "-c",
                        "-t", "3", "-s", cfg], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0
    assert re.search(r"^GPU ysum = \d+\.\d{20}$", r.stdout, re.M)
    rel = float(re.search(r"ysum relative difference: ([0-9.e-]+)", r.stdout).group(1))
    assert rel <= 1e-8


def test_sgemm_app_usage_errors():
    r = subprocess.run([app("gemm/hml_sgemm"), "-M", "10"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "cannot be zero" in r.stderr
    r = subprocess.run([app("gemm/hml_sgemm"), "-Q"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "Unknown option" in r.stderr
