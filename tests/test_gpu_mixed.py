"""Mixed precision (FMMH_OPT_PRECISION = 1): the M2M, M2L and L2L translations
run on FP32 MFMA with FP64 sums into the expansions; P2M, L2P and the near
field stay FP64.  Checks that the FP32 paths are the ones that run (results
differ from FP64 at the FP32 level, not at rounding level) and that their
error stays far below the truncation error the FMM already has against the
direct sum."""
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


@pytest.mark.parametrize("p", [2, 6, 10])
def test_mixed_per_target_mfma_for_many_sets(mixed, p):
    """Many sets: the per-target MFMA M2L in FP32 (FP64 sum over pairs) differs
    from the FP64 form at the FP32 level only, for every set, including sets
    beyond the first group of 16."""
    g = load_golden("water999_dipole_orders.npz")
    x, e = g["coordinates"], g["exclusions"]
    n = x.shape[0]
    rng = np.random.default_rng(p)
    S = 20
    m = np.zeros((n, 4, S), order="F")
    for k in range(S):
        m[:, 0, k] = g["multipoles"][:, 0]
        m[:, 1:4, k] = rng.normal(0.0, 0.05, (n, 3))
    ref = mixed.field_fmm_sets(0, x, m, e, 0.5, 64, p, 1)
    mixed.set_precision(True)
    got = mixed.field_fmm_sets(0, x, m, e, 0.5, 64, p, 1)
    for k in (0, 15, 16, S - 1):
        d = max(block_rel(got[:, :, k], ref[:, :, k], 1))
        assert 1e-12 < d < 2e-6, (k, d)


@pytest.mark.parametrize("p,nsets", [(3, 8), (6, 8), (10, 8), (6, 20)])
def test_mixed_m2m_moments_close_to_fp64(mixed, p, nsets):
    """The FP32 M2M (box-scaled shift matrices per octant) against the FP64
    one on a 30k-site water box: leaf moments (P2M, FP64 in both) are
    identical; the levels with few (node, set) columns keep the FP64 kernel
    (identical too); every other internal node agrees to FP32 level relative
    to the node's scaled norm, and some do differ (the FP32 form ran)."""
    from fmm_amd import synth
    x, q, e, _, rng = synth.water_box(10000, 7 + p)
    n = x.shape[0]
    m = np.zeros((n, 4, nsets), order="F")
    for k in range(nsets):
        m[:, 0, k] = q
        m[:, 1:4, k] = rng.normal(0.0, 0.05, (n, 3))
    tr = mixed.tree(x, 64, 0.5)
    nn = len(tr["r"])
    nm = (p + 1) * (p + 2) * (p + 3) // 6
    deg = np.concatenate([[k] * ((k + 1) * (k + 2) // 2) for k in range(p + 1)])
    mixed.field_fmm_sets(0, x, m, e, 0.5, 64, p, 1)
    c64 = mixed.debug_buffer("e_cm", nn * nm * nsets).reshape(nsets, nn, nm)
    mixed.set_precision(True)
    mixed.field_fmm_sets(0, x, m, e, 0.5, 64, p, 1)
    c32 = mixed.debug_buffer("e_cm", nn * nm * nsets).reshape(nsets, nn, nm)
    leaf = tr["is_leaf"] != 0
    np.testing.assert_array_equal(c32[:, leaf], c64[:, leaf])
    scale = tr["r"][:, None] ** (-deg[None, :])            # box-scaled moments M(s) r^-|s|
    worst = 0.0
    for k in range(nsets):
        a, b = c32[k][~leaf] * scale[~leaf], c64[k][~leaf] * scale[~leaf]
        err = np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)
        assert err.max() < 2e-6, (k, err.max())
        worst = max(worst, err.max())
    assert worst > 1e-12
