"""Many multipole sets on one tree (fmmh_field_*_sets): every set must give
what the single-set call gives for that set.  Every call goes through the C
ABI of libfmm_hip.so.

Bars: bit-identical where the set path runs the same arithmetic as the
single-set path (direct sums, VALU M2L); <= 1e-12 relative per output-order
block where the M2L runs on the FP64 matrix cores (another summation order);
linearity in the multipoles at the full water size."""
import numpy as np
import pytest

import oracle
from conftest import block_rel, load_golden, rel
from fmm_amd import synth

pytestmark = pytest.mark.gpu
PARITY = 1e-12


def _sets(n, dim, nsets, seed):
    rng = np.random.default_rng(seed)
    m = np.stack([synth.random_multipoles(rng, n, dim) for _ in range(nsets)], axis=2)
    return np.asfortranarray(m)


def _single(ctx, x, m, e, F, damp, df, p=None):
    out = []
    for k in range(m.shape[2]):
        mk = np.asfortranarray(m[:, :, k])
        if p is None:
            out.append(ctx.field_direct(0, x, mk, e, F, damp, df))
        else:
            out.append(ctx.field_fmm(0, x, mk, e, 0.5, 64, p, F, damp, df))
    return np.stack(out, axis=2)


@pytest.fixture
def mfma_mode(gpu_ctx):
    yield gpu_ctx.set_m2l_mfma
    gpu_ctx.set_m2l_mfma(1)


@pytest.mark.parametrize("dim,F,damp", [(1, 1, ""), (4, 2, "THOLE"), (10, 1, "AMOEBA"), (10, 2, "ERF")])
def test_direct_sets_bit_identical(gpu_ctx, dim, F, damp):
    g = load_golden("water999_q.npz")
    x, e = g["coordinates"], g["exclusions"]
    df = g["erf_df"] if damp == "ERF" else g["thole_df"]
    m = _sets(x.shape[0], dim, 5, 11)
    got = gpu_ctx.field_direct_sets(0, x, m, e, F, damp, df)
    np.testing.assert_array_equal(got, _single(gpu_ctx, x, m, e, F, damp, df))


@pytest.mark.parametrize("nsets", [2, 3, 5, 7])
def test_fmm_sets_few_match_single_set(gpu_ctx, mfma_mode, nsets):
    """Below 6 sets the M2L runs set by set (class-grouped GEMM): the same
    arithmetic as the single-set call, so the fields are identical bit for
    bit.  From 6 sets on the per-target MFMA form takes over (one padded
    16-set tile instead of 6+ GEMM passes): rounding-level agreement."""
    g = load_golden("water999_q.npz")
    x, e, df = g["coordinates"], g["exclusions"], g["thole_df"]
    m = _sets(x.shape[0], 4, nsets, 12)
    got = gpu_ctx.field_fmm_sets(0, x, m, e, 0.5, 64, 6, 2, "THOLE", df)
    ref = _single(gpu_ctx, x, m, e, 2, "THOLE", df, p=6)
    if nsets < 6:
        np.testing.assert_array_equal(got, ref)
    else:
        assert not np.array_equal(got, ref)   # the MFMA form ran
        for k in range(nsets):
            assert max(block_rel(got[:, :, k], ref[:, :, k], 2)) < 1e-13


@pytest.mark.parametrize("p", list(range(0, 11)))
def test_m2l_mfma_matches_valu_every_order(gpu_ctx, mfma_mode, p):
    """MFMA M2L forced on for 3 sets (zero-padded to 16) against the VALU
    M2L, every order the MFMA form supports: checks the fragment layout."""
    g = load_golden("water999_dipole_orders.npz")
    x, e = g["coordinates"], g["exclusions"]
    m = _sets(x.shape[0], 4, 3, 13)
    mfma_mode(0)
    ref = gpu_ctx.field_fmm_sets(0, x, m, e, 0.5, 64, p, 1)
    mfma_mode(2)
    got = gpu_ctx.field_fmm_sets(0, x, m, e, 0.5, 64, p, 1)
    for k in range(3):
        assert max(block_rel(got[:, :, k], ref[:, :, k], 1)) < PARITY


@pytest.mark.parametrize("p", list(range(0, 11)))
def test_m2l_class_grouped_matches_valu_every_order(gpu_ctx, mfma_mode, p):
    """Class-grouped FP64 MFMA GEMM M2L (the default for one set) against the
    per-target VALU contraction, every order 0..10."""
    g = load_golden("water999_dipole_orders.npz")
    x, m, e = g["coordinates"], g["multipoles"], g["exclusions"]
    mfma_mode(0)
    ref = gpu_ctx.field_fmm(0, x, m, e, 0.5, 64, p, 1)
    mfma_mode(1)
    got = gpu_ctx.field_fmm(0, x, m, e, 0.5, 64, p, 1)
    assert max(block_rel(got, ref, 1)) < PARITY


def test_m2l_class_grouped_for_many_sets(gpu_ctx, mfma_mode):
    """Mode 3 forces the class-grouped GEMM for 18 sets (set by set)."""
    g = load_golden("water999_q.npz")
    x, e = g["coordinates"], g["exclusions"]
    m = _sets(x.shape[0], 4, 18, 21)
    ref = gpu_ctx.field_fmm_sets(0, x, m, e, 0.5, 64, 6, 1)
    mfma_mode(3)
    got = gpu_ctx.field_fmm_sets(0, x, m, e, 0.5, 64, 6, 1)
    for k in range(18):
        assert max(block_rel(got[:, :, k], ref[:, :, k], 1)) < PARITY, k


@pytest.mark.parametrize("nsets,p", [(16, 6), (20, 5), (32, 6), (40, 8)])
def test_fmm_sets_mfma_vs_single_calls(gpu_ctx, nsets, p):
    """From 16 sets on the M2L runs on MFMA in chunks of up to 32 sets (40 =
    32 + 8, the second chunk zero-padded to 16)."""
    g = load_golden("water999_q.npz")
    x, e, df = g["coordinates"], g["exclusions"], g["thole_df"]
    m = _sets(x.shape[0], 10, nsets, 14)
    got = gpu_ctx.field_fmm_sets(0, x, m, e, 0.5, 64, p, 1, "AMOEBA", df)
    ref = _single(gpu_ctx, x, m, e, 1, "AMOEBA", df, p=p)
    for k in range(nsets):
        assert max(block_rel(got[:, :, k], ref[:, :, k], 1)) < PARITY, k


def test_fmm_sets_vs_oracle(gpu_ctx):
    g = load_golden("water999_q.npz")
    x, e, df = g["coordinates"], g["exclusions"], g["thole_df"]
    m = _sets(x.shape[0], 4, 17, 15)
    got = gpu_ctx.field_fmm_sets(0, x, m, e, 0.5, 64, 6, 2, "THOLE", df)
    for k in (0, 9, 16):
        ref = oracle.field_fmm(x, np.asfortranarray(m[:, :, k]), e, 0.5, 64, 6, 2, "THOLE", df)
        assert max(block_rel(got[:, :, k], ref, 2)) < PARITY, k


def test_fmm_sets_linear_at_full_size(gpu_ctx):
    """Size-independent check at the 200k-site water size: the field is
    linear in the multipoles, so set 2 = set 0 + 3 * set 1 holds to rounding
    (MFMA path: 20 sets)."""
    cfg = synth.config(2)
    x, e, df = cfg["coordinates"], cfg["exclusions"], cfg["damping_factors"]
    m = np.array(cfg["multipoles"][:, :, :20], order="F")
    m[:, :, 2] = m[:, :, 0] + 3.0 * m[:, :, 1]
    got = gpu_ctx.field_fmm_sets(0, x, m, e, 0.5, 64, 6, 1, "AMOEBA", df)
    comb = got[:, :, 0] + 3.0 * got[:, :, 1]
    assert rel(got[:, :, 2], comb) < 1e-11
    # and set 5 alone through the single-set path
    one = gpu_ctx.field_fmm(0, x, np.asfortranarray(m[:, :, 5]), e, 0.5, 64, 6, 1, "AMOEBA", df)
    assert max(block_rel(got[:, :, 5], one, 1)) < PARITY


def test_sets_device_entry_points_match_host(gpu_ctx):
    import torch
    g = load_golden("water999_q.npz")
    dev = torch.device("cuda:0")
    x, e, df = g["coordinates"], g["exclusions"], g["thole_df"]
    m = _sets(x.shape[0], 4, 18, 16)

    def col(a, dt=torch.float64):
        a = np.asarray(a)
        perm = tuple(range(a.ndim))[::-1]
        return torch.as_tensor(np.ascontiguousarray(a.transpose(perm)), dtype=dt, device=dev).permute(perm)

    X, M, E = col(x), col(m), col(e, torch.int32)
    D = torch.as_tensor(df, device=dev)
    out = col(np.zeros((x.shape[0], 10, 18)))
    gpu_ctx.field_fmm_sets_device(X, M, E, 0.5, 64, 6, 2, out, "THOLE", D)
    np.testing.assert_array_equal(out.cpu().numpy(), gpu_ctx.field_fmm_sets(0, x, m, e, 0.5, 64, 6, 2, "THOLE", df))
    out.zero_()
    gpu_ctx.field_direct_sets_device(X, M, E, 2, out, "THOLE", D)
    np.testing.assert_array_equal(out.cpu().numpy(), gpu_ctx.field_direct_sets(0, x, m, e, 2, "THOLE", df))


def test_sets_field_is_accumulated(gpu_ctx):
    g = load_golden("water999_q.npz")
    x, e = g["coordinates"], g["exclusions"]
    m = _sets(x.shape[0], 4, 3, 17)
    base = np.asfortranarray(np.random.default_rng(3).normal(size=(x.shape[0], 4, 3)))
    got = gpu_ctx.field_fmm_sets(0, x, m, e, 0.5, 64, 5, 1, field=base.copy(order="F"))
    fresh = gpu_ctx.field_fmm_sets(0, x, m, e, 0.5, 64, 5, 1)
    np.testing.assert_allclose(got, base + fresh, rtol=0, atol=1e-13 * np.abs(fresh).max())


def test_sets_errors(gpu_ctx):
    from fmm_amd.octree import FmmError
    g = load_golden("water999_q.npz")
    x, e = g["coordinates"], g["exclusions"]
    m = np.zeros((x.shape[0], 4, 0), order="F")
    with pytest.raises(FmmError, match="nsets"):
        gpu_ctx.field_fmm_sets(0, x, m, e, 0.5, 64, 5, 1)
    with pytest.raises(FmmError, match="mode"):
        gpu_ctx.set_m2l_mfma(5)


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("nsets", [16, 20, 32])
def test_autotuned_tile_shape_gives_identical_results(nsets, mixed):
    """The first many-sets call times every valid tile-shape variant of the
    per-target MFMA M2L and keeps the fastest; every variant computes the same
    sums in the same order, so the tuned result equals the untuned default bit
    for bit, and later calls reuse the choice."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from fmm_amd.octree import Context
    g = load_golden("water999_dipole_orders.npz")
    x, e = g["coordinates"], g["exclusions"]
    m = _sets(x.shape[0], 4, nsets, 29)
    ctx = Context(0)
    try:
        ctx.set_precision(mixed)
        ctx.set_autotune(False)
        ref = ctx.field_fmm_sets(0, x, m, e, 0.5, 64, 7, 1)
        ctx.set_autotune(True)
        tuned = ctx.field_fmm_sets(0, x, m, e, 0.5, 64, 7, 1)
        choices, ms = ctx.tuning()
        assert len(choices) == 1 and (ms >= 0).sum() >= 1
        again = ctx.field_fmm_sets(0, x, m, e, 0.5, 64, 7, 1)
        assert ctx.tuning()[0] == choices
        assert np.array_equal(tuned, ref) and np.array_equal(again, ref)
    finally:
        ctx.close()
