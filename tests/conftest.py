import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# Selection tests expect the table / heuristic selectors: keep any learned
# selector named in the environment out of this process
# (tests/test_gpu_mlp.py checks loading in child processes of its own).
os.environ["HML_SELECTOR"] = "off"
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


@pytest.fixture(scope="session")
def hml():
    """The product library; GPU tests fail (not skip) if it cannot load or
    no device is visible."""
    import hiperml_amd
    assert hiperml_amd.device_count() > 0, "no GPU visible to libhml.so"
    hiperml_amd.check(hiperml_amd.lib.hmlInit())
    return hiperml_amd
