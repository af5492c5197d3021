"""Shared test setup: import paths, the `gpu` marker, and a GPU context."""
import os
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(REPO / "impatient-mri_amd"))
sys.path.insert(0, str(REPO))
GOLDEN = REPO / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (run with -m gpu)")


@pytest.fixture(scope="session")
def ctx():
    from impatient import Context
    c = Context(int(os.environ.get("IMP_TEST_DEVICE", "0")))
    yield c
    c.close()


def rel_l2(a, b):
    import numpy as np
    a = np.asarray(a, np.complex128).reshape(-1)
    b = np.asarray(b, np.complex128).reshape(-1)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
