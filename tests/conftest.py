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


def ncomp(k):
    return (k + 1) * (k + 2) * (k + 3) // 6


def block_rel(a, b, field_order, skip_potential=False):
    """Relative error per output-order block: components of degree k are columns ncomp(k-1)..ncomp(k)-1."""
    out = []
    for k in range(field_order + 1):
        if skip_potential and k == 0 and field_order > 0:
            continue
        lo, hi = ncomp(k - 1), ncomp(k)
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
