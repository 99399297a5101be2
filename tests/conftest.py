import os
import sys

import numpy as np
import pytest

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
