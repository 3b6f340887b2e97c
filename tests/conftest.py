import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libfmm_hip.so")


def rel(a, b):
    """Relative 2-norm error ||a - b|| / ||b||."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


BLOCKS = [(0, 1), (1, 4), (4, 10), (10, 20)]


def block_rel(a, b, field_order, skip_potential=False):
    out = []
    for k, (lo, hi) in enumerate(BLOCKS[: field_order + 1]):
        if skip_potential and k == 0:
            continue
        out.append(rel(a[:, lo:hi], b[:, lo:hi]))
    return out


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from fmm_amd.octree import Context
    ctx = Context(0)
    yield ctx
    ctx.close()
