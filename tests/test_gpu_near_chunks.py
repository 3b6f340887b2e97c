"""Near-field work split (FMMH_OPT_NEAR_CHUNK): a leaf whose pair work exceeds
a cap runs as several workgroups over consecutive parts of its source list,
and their partial fields are summed in chunk order.  A small cap forces the
split on the 999-site fixtures (the automatic cap splits only at larger
sizes, e.g. the 200k-site water box of the linearity and bench checks).

Bars: <= 1e-12 relative per output-order block against the reference's own
output and the oracle; bit-identical between the sets path and single-set
calls under the same split (the split depends on the tree only).  Every call
goes through the C ABI of libfmm_hip.so."""
import numpy as np
import pytest

import oracle
from conftest import block_rel, load_golden

pytestmark = pytest.mark.gpu
PARITY = 1e-12
CAP = 24


@pytest.fixture
def chunked(gpu_ctx):
    gpu_ctx.set_near_chunk(CAP)
    yield gpu_ctx
    gpu_ctx.set_near_chunk(0)


@pytest.mark.parametrize("damp", ["", "THOLE", "AMOEBA", "ERF"])
def test_split_matches_reference(chunked, damp):
    g = load_golden("water999_q.npz")
    df = g["erf_df"] if damp == "ERF" else g["thole_df"]
    key = damp.lower() or "none"
    ff = chunked.field_fmm(0, g["coordinates"], g["multipoles"], g["exclusions"], 0.5, 64, 6, 2, damp, df)
    leaves = chunked.stats()["num_leaves"]
    assert chunked.stats()["near_workgroups"] > 2 * leaves   # the split ran
    assert max(block_rel(ff, g[f"fmm_{key}"], 2)) < PARITY


def test_split_matches_unsplit_and_oracle(chunked):
    g = load_golden("water999_dipole_orders.npz")
    x, m, e = g["coordinates"], g["multipoles"], g["exclusions"]
    got = chunked.field_fmm(0, x, m, e, 0.5, 64, 5, 1)
    chunked.set_near_chunk(1e12)          # no leaf reaches this cap: one workgroup per leaf
    whole = chunked.field_fmm(0, x, m, e, 0.5, 64, 5, 1)
    assert chunked.stats()["near_workgroups"] == chunked.stats()["num_leaves"]
    assert not np.array_equal(got, whole)              # another summation order
    assert max(block_rel(got, whole, 1)) < 1e-13
    ref = oracle.field_fmm(x, m, e, 0.5, 64, 5, 1)
    assert max(block_rel(got, ref, 1)) < PARITY


@pytest.mark.parametrize("nsets", [3, 8])
def test_split_sets_bit_identical_to_single_calls(chunked, nsets):
    g = load_golden("water999_q.npz")
    x, e, df = g["coordinates"], g["exclusions"], g["thole_df"]
    rng = np.random.default_rng(51)
    m = np.asfortranarray(rng.normal(size=(x.shape[0], 4, nsets)))
    chunked.set_m2l_mfma(0)          # VALU M2L: the same arithmetic for one set and many
    try:
        got = chunked.field_fmm_sets(0, x, m, e, 0.5, 64, 6, 2, "THOLE", df)
        for k in range(nsets):
            one = chunked.field_fmm(0, x, np.asfortranarray(m[:, :, k]), e, 0.5, 64, 6, 2, "THOLE", df)
            np.testing.assert_array_equal(got[:, :, k], one)
    finally:
        chunked.set_m2l_mfma(1)


def test_split_with_periodic_images(chunked):
    g = load_golden("water999_periodic.npz")
    x, m, e, box = g["coordinates"], g["multipoles"], g["exclusions"], g["box_cube"]
    got = chunked.field_fmm(0, x, m, e, 0.5, 64, 6, 2, "THOLE", g["thole_df"], is_periodic=True, box_matrix=box)
    assert max(block_rel(got, g["cube_thole_p6_f2"], 2)) < PARITY


def test_split_with_tree_reuse_and_graphs():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from fmm_amd.octree import Context
    g = load_golden("water999_q.npz")
    x, m, e, df = g["coordinates"], g["multipoles"], g["exclusions"], g["thole_df"]
    c = Context(0)
    try:
        c.set_near_chunk(CAP)
        first = c.field_fmm(0, x, m, e, 0.5, 64, 6, 2, "THOLE", df)
        c.set_tree_reuse(True)
        c.set_graph(True)
        for _ in range(4):
            np.testing.assert_array_equal(c.field_fmm(0, x, m, e, 0.5, 64, 6, 2, "THOLE", df), first)
        assert c.stats()["graph"] == 2
    finally:
        c.close()
