"""Periodic field_fmm (is_periodic = .true.) against the reference's own
output on the same inputs (tests/golden/water999_periodic.npz, written by the
compiled reference): 27 image lists, the lattice sum beyond them and the
dipole correction.  The fixtures are boxes where the reference's lattice sum
converges (an exact cube; a cell without net dipole in a triclinic box);
elsewhere its output is rounding noise.  Every call goes through the C ABI
of libfmm_hip.so."""
import numpy as np
import pytest

from conftest import block_rel, load_golden

pytestmark = pytest.mark.gpu
PARITY = 1e-12
CASES = ["cube_thole_p6_f2", "cube_none_p4_f1", "cube_amoeba_p5_f1", "cube_erf_p5_f2", "sym_triclinic_none_p5_f1",
         "sym_triclinic_thole_p6_f2"]


def _args(g, key):
    sym = key.startswith("sym_")
    pre = "sym_" if sym else ""
    parts = key.split("_")[1:] if sym else key.split("_")
    box = g["box_triclinic"] if sym else g["box_cube"]
    damp = {"thole": "THOLE", "none": "", "amoeba": "AMOEBA", "erf": "ERF"}[parts[1]]
    p, F = int(parts[2][1:]), int(parts[3][1:])
    df = g[pre + ("erf_df" if damp == "ERF" else "thole_df")]
    return g[pre + "coordinates"], g[pre + "multipoles"], g[pre + "exclusions"], box, damp, df, p, F


@pytest.mark.parametrize("key", CASES)
def test_periodic_vs_reference(gpu_ctx, key):
    g = load_golden("water999_periodic.npz")
    x, m, e, box, damp, df, p, F = _args(g, key)
    got = gpu_ctx.field_fmm(0, x, m, e, 0.5, 64, p, F, damp, df, is_periodic=True, box_matrix=box)
    assert max(block_rel(got, g[key], F, damp == "AMOEBA")) < PARITY
    assert int(gpu_ctx.debug_buffer("p_levels", 1, np.int32)[0]) == int(g[key + "_levels"])


@pytest.mark.parametrize("key", ["cube_thole_p6_f2", "sym_triclinic_none_p5_f1"])
def test_periodic_mixed_precision_close(gpu_ctx, key):
    g = load_golden("water999_periodic.npz")
    x, m, e, box, damp, df, p, F = _args(g, key)
    gpu_ctx.set_precision(True)
    try:
        got = gpu_ctx.field_fmm(0, x, m, e, 0.5, 64, p, F, damp, df, is_periodic=True, box_matrix=box)
    finally:
        gpu_ctx.set_precision(False)
    assert max(block_rel(got, g[key], F, True)) < 2e-6


def test_periodic_sets_and_reuse(gpu_ctx):
    g = load_golden("water999_periodic.npz")
    x, m0, e, box, damp, df, p, F = _args(g, "sym_triclinic_thole_p6_f2")
    rng = np.random.default_rng(8)
    m = np.asfortranarray(np.stack([m0, rng.normal(0, 0.3, m0.shape)], axis=2))
    sets = gpu_ctx.field_fmm_sets(0, x, m, e, 0.5, 64, p, F, damp, df, is_periodic=True, box_matrix=box)
    one = gpu_ctx.field_fmm(0, x, np.asfortranarray(m[:, :, 0]), e, 0.5, 64, p, F, damp, df, is_periodic=True,
                            box_matrix=box)
    np.testing.assert_array_equal(sets[:, :, 0], one)
    assert max(block_rel(sets[:, :, 0], g["sym_triclinic_thole_p6_f2"], F)) < PARITY
    gpu_ctx.set_tree_reuse(True)
    gpu_ctx.set_timing(True)
    try:
        a = gpu_ctx.field_fmm(0, x, m0, e, 0.5, 64, p, F, damp, df, is_periodic=True, box_matrix=box)
        b = gpu_ctx.field_fmm(0, x, m0, e, 0.5, 64, p, F, damp, df, is_periodic=True, box_matrix=box)
        assert "tree_reused" in gpu_ctx.phase_times()
        np.testing.assert_array_equal(a, b)
        # a different box must rebuild the image lists; no box at all as well
        box2 = np.array(box, order="F")
        box2[0, 1] = 2.5
        gpu_ctx.field_fmm(0, x, m0, e, 0.5, 64, p, F, damp, df, is_periodic=True, box_matrix=box2)
        assert "lists" in gpu_ctx.phase_times()
        gpu_ctx.field_fmm(0, x, m0, e, 0.5, 64, p, F, damp, df)
        assert "lists" in gpu_ctx.phase_times()
    finally:
        gpu_ctx.set_tree_reuse(False)
        gpu_ctx.set_timing(False)


def test_periodic_errors(gpu_ctx):
    from fmm_amd.octree import FmmError
    g = load_golden("water999_periodic.npz")
    x, m, e = g["coordinates"], g["multipoles"], g["exclusions"]
    with pytest.raises(FmmError, match="Not enough coordinates for periodic FMM"):
        gpu_ctx.field_fmm(0, x[:50], m[:50], np.asfortranarray(e[:50, :1]), 0.5, 64, 4, 1, is_periodic=True,
                          box_matrix=g["box_cube"])
    with pytest.raises(FmmError, match="expansion_order >= 2"):
        gpu_ctx.field_fmm(0, x, m, e, 0.5, 64, 1, 1, is_periodic=True, box_matrix=g["box_cube"])
