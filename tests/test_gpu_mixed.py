"""Mixed precision (FMMH_OPT_PRECISION = 1): the M2L translations run on FP32
MFMA with an FP64 sum into the local expansions; everything else is FP64.
Checks that the FP32 path is the one that runs (results differ from FP64 at
the FP32 level, not at rounding level) and that its error stays far below
the truncation error the FMM already has against the direct sum."""
import numpy as np
import pytest

from conftest import block_rel, load_golden, rel

pytestmark = pytest.mark.gpu


@pytest.fixture
def mixed(gpu_ctx):
    yield gpu_ctx
    gpu_ctx.set_precision(False)


@pytest.mark.parametrize("p", [4, 6, 8, 10])
def test_mixed_m2l_close_to_fp64(mixed, p):
    g = load_golden("water999_dipole_orders.npz")
    x, m, e = g["coordinates"], g["multipoles"], g["exclusions"]
    ref = mixed.field_fmm(0, x, m, e, 0.5, 64, p, 1)
    mixed.set_precision(True)
    got = mixed.field_fmm(0, x, m, e, 0.5, 64, p, 1)
    d = max(block_rel(got, ref, 1))
    assert 1e-12 < d < 2e-6, d


def test_mixed_error_budget_vs_direct(mixed):
    """Truncation (FP64 FMM vs direct) dominates: mixed adds < 1 % to it."""
    g = load_golden("uniform4096_p5.npz")
    x, m, e = g["coordinates"], g["multipoles"], g["exclusions"]
    direct = g["field_direct"]
    f64 = mixed.field_fmm(0, x, m, e, 0.5, 64, 8, 1)
    mixed.set_precision(True)
    f32 = mixed.field_fmm(0, x, m, e, 0.5, 64, 8, 1)
    t64 = rel(f64[:, 1:4], direct[:, 1:4])
    t32 = rel(f32[:, 1:4], direct[:, 1:4])
    assert t64 < 1e-4
    assert abs(t32 - t64) < 0.01 * t64 + 1e-7, (t32, t64)


def test_mixed_option_errors(gpu_ctx):
    assert gpu_ctx._lib.fmmh_set_option(gpu_ctx.handle, 5, 2.0) == 1
    assert b"precision" in gpu_ctx._lib.fmmh_last_error()
