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
    # Some GPU tests check against FP64 restatements run with torch on the
    # GPU (tests/fp64_torch.py).  torch's HIP runtime has to be initialised
    # before libimpatient.so is loaded: once the library has brought in the
    # system HIP runtime, torch finds no device.
    try:
        import torch
        torch.cuda.is_available()
    except ImportError:
        pass
    from impatient import Context
    c = Context(int(os.environ.get("IMP_TEST_DEVICE", "0")))
    yield c
    c.close()


def rel_l2(a, b):
    import numpy as np
    a = np.asarray(a, np.complex128).reshape(-1)
    b = np.asarray(b, np.complex128).reshape(-1)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def difference_coo(shape):
    """[D_0; D_1; ...]: one block of periodic first differences
    (Dp)(i) = p(i) - p(i + e_b) per image axis b, 0-based COO."""
    import numpy as np
    import scipy.sparse as sp
    n = int(np.prod(shape))
    blocks = []
    for b, na in enumerate(shape):
        D = sp.identity(na, format="csr") - sp.csr_matrix(
            (np.ones(na), (np.arange(na), (np.arange(na) + 1) % na)), shape=(na, na))
        K = None
        for a, s in enumerate(shape):
            m = D if a == b else sp.identity(s, format="csr")
            K = m if K is None else sp.kron(K, m, format="csr")
        blocks.append(K)
    C = sp.vstack(blocks).tocoo()
    return C.shape[0], n, C.row.astype(np.int32), C.col.astype(np.int32), C.data
