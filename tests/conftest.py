import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O  # noqa: the CPU oracle (test infrastructure)
    O.lib()
    return O


def rel_err(a, b, scale=None):
    import numpy as np
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    den = np.max(np.abs(b)) if scale is None else scale
    return float(np.max(np.abs(a - b)) / max(den, 1e-300))
