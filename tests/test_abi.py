"""The C-ABI library loads and exports, unmangled, every function that
include/*.h declares (no compute calls: runs without a GPU)."""
import ctypes
import glob
import os
import re
import subprocess

import hiperml_amd as hml
from hiperml_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b(hml[A-Z]\w*)\s*\(", re.M)


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(DECL.findall(text))
    return names


def test_headers_declare_reference_api():
    names = declared()
    for n in ("hmlSgemm", "hmlSgemmBasic", "hmlSgemmKernelConfig", "hmlSgemv",
              "hmlSgemvBasic", "hmlSgemvKernelConfig"):
        assert n in names


def test_every_declared_symbol_exported():
    names = declared()
    assert len(names) > 30
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = sorted(n for n in names if n not in exported)
    assert not missing, missing
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for n in names:
        getattr(lib, n)


def test_python_binding_covers_header():
    names = declared()
    assert names <= set(_lib.EXPORTED), sorted(names - set(_lib.EXPORTED))


def test_status_strings():
    assert hml.lib.hmlStatusString(0) == b"success"
    assert hml.lib.hmlStatusString(7) == b"workspace too small"


def test_argument_errors_without_gpu():
    # rejected before any device work
    assert hml.lib.hmlSgemmEx(0, 0, 0, -1, 4, 4, 1.0, None, 4, None, 4, 0.0,
                              None, 4, 0, None, 0, None) == 1
    assert b"negative" in hml.lib.hmlLastError()
    assert hml.lib.hmlSgemmEx(0, 0, 0, 16, 4, 8, 1.0, 1, 4, 1, 4, 0.0, 1, 4,
                              0, None, 0, None) == 1          # lda < K
    assert hml.lib.hmlSgemmEx(2, 0, 0, 16, 4, 8, 1.0, 1, 8, 1, 4, 0.0, 1, 4,
                              0, None, 0, None) == 1          # bad order
    assert hml.lib.hmlSgemvEx(0, 0, 16, 8, 1.0, 1, 8, 1, 2, 0.0, 1, 1, 0,
                              None, 0, None) == 6             # incx != 1
    # empty problems succeed without touching memory
    assert hml.lib.hmlSgemmEx(0, 0, 0, 0, 4, 4, 1.0, None, 4, None, 4, 0.0,
                              None, 4, 0, None, 0, None) == 0
    # k-means assignment: dims 1..64, clusters >= 1, known flags, pointers
    km = hml.lib.hmlKmeansAssign
    assert km(1, 1, 10, 0, 1, 4, 0, None) == 1
    assert km(1, 1, 10, 65, 1, 4, 0, None) == 1
    assert b"1..64" in hml.lib.hmlLastError()
    assert km(1, 1, 10, 8, 1, 0, 0, None) == 1
    assert km(1, 1, 10, 8, 1, 4, 8, None) == 1         # unknown flag
    assert km(1, 1, 10, 8, 1, 4, 6, None) == 1         # both kernels forced
    assert km(None, 1, 10, 8, 1, 4, 0, None) == 1
    assert km(None, None, 0, 8, None, 4, 1, None) == 0


def test_no_torch_or_oracle_in_product():
    # the product never imports the oracle or a tensor library
    for py in glob.glob(os.path.join(ROOT, "hiperml_amd", "*.py")):
        src = open(py).read()
        assert "oracle" not in src.replace("no CPU fallback", "")
        assert "import torch" not in src
