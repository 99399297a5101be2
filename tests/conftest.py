import os
import sys

import numpy as np
import pytest

# torch bundles its own HIP runtime under the same soname (libamdhip64.so.7)
# as the system one libprimus_hip.so links against; whichever loads first
# serves the process, and torch only initialises its devices on its own.  So
# torch is loaded before any test loads the library (the graph tests use it).
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN_PATH = os.path.join(ROOT, "tests", "golden", "ntt_golden.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


def load_golden():
    """Golden vectors written by tests/golden/make_golden.py: {label: {field: array}}."""
    z = np.load(GOLDEN_PATH, allow_pickle=False)
    cases = {}
    for key in z.files:
        label, field = key.split("/")
        cases.setdefault(label, {})[field] = z[key]
    return cases


@pytest.fixture(scope="session")
def golden():
    return load_golden()
