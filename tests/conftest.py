import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "senga-versiona_amd"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests", "golden"), os.path.join(ROOT, "senga-versiona_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


@pytest.fixture(scope="session")
def oracle_built():
    """The oracle shared library (built here if missing)."""
    import build as _build
    path = os.path.join(ROOT, "oracle", "lib", "liboracle.so")
    if not os.path.exists(path):
        _build.build_oracle()
    return path


def rel_err(a, b):
    """max|a-b| / max|b| (max-norm relative error of a field)."""
    import numpy as np
    den = float(np.max(np.abs(b)))
    return float(np.max(np.abs(a - b))) / (den if den > 0 else 1.0)
