"""QM/MM embedding routines on the GPU (fmmh_get_boxes, fmmh_get_multipoles,
fmmh_interpolate_electron_fields) against the reference's own
interpolate_electron_fields output and the pinned restatement (the
reference's get_boxes / get_multipoles abort before returning).  Every call
goes through the C ABI of libfmm_hip.so."""
import numpy as np
import pytest

import oracle
from conftest import block_rel, load_golden

pytestmark = pytest.mark.gpu
CASES = [("t5_n64_p4", 0.5, 64, 4), ("t3_n16_p6", 0.3, 16, 6), ("t5_n32_p5", 0.5, 32, 5)]


@pytest.mark.parametrize("key,theta,ncrit,p", CASES)
def test_get_boxes_exact(gpu_ctx, key, theta, ncrit, p):
    g = load_golden("water999_qmmm.npz")
    x, tg = g["coordinates"], g["targets"]
    near, bc, bi = gpu_ctx.get_boxes(x, tg, theta, ncrit, p)
    rn, rc, ri = oracle.get_boxes(x, tg, theta, ncrit, p, impl="oracle")
    np.testing.assert_array_equal(near, rn)
    np.testing.assert_array_equal(bi, ri)
    np.testing.assert_array_equal(bc, rc)
    np.testing.assert_array_equal(bi, g[key + "_box_indices"])


@pytest.mark.parametrize("key,theta,ncrit,p", CASES)
def test_get_multipoles(gpu_ctx, key, theta, ncrit, p):
    g = load_golden("water999_qmmm.npz")
    x, m2, tg = g["coordinates"], g["multipoles"], g["targets"]
    near, bc, bm = gpu_ctx.get_multipoles(0, x, m2, tg, theta, ncrit, p)
    rn, rc, rm = oracle.get_multipoles(x, m2, tg, theta, ncrit, p, impl="oracle")
    np.testing.assert_array_equal(near, rn)
    np.testing.assert_array_equal(bc, rc)
    assert bm.shape == rm.shape
    for d in range(bm.shape[2]):
        num = np.linalg.norm(bm[:, :, d] - rm[:, :, d], axis=1)
        assert np.all(num <= 1e-12 * np.linalg.norm(rm[:, :, d], axis=1))


@pytest.mark.parametrize("key,theta,ncrit,p", CASES)
def test_interpolate_vs_reference(gpu_ctx, key, theta, ncrit, p):
    g = load_golden("water999_qmmm.npz")
    F = int(g[key + "_field_order"])
    got = gpu_ctx.interpolate_electron_fields(0, g["coordinates"], g[key + "_local_expansion"],
                                              g[key + "_box_indices"], theta, ncrit, p, F)
    for d in range(got.shape[2]):
        assert max(block_rel(got[:, :, d], g[key + "_field"][:, :, d], F)) < 1e-12


def test_qmmm_errors(gpu_ctx):
    from fmm_amd.octree import Context, FmmError
    g = load_golden("water999_qmmm.npz")
    x = g["coordinates"]
    le = g["t5_n64_p4_local_expansion"]
    bad = np.array(g["t5_n64_p4_box_indices"])
    bad[0] = 10 ** 6
    with pytest.raises(FmmError, match="box_indices"):
        gpu_ctx.interpolate_electron_fields(0, x, le, bad, 0.5, 64, 4, 1)
    fresh = Context(0)
    try:
        assert fresh._lib.fmmh_get_boxes_export(fresh.handle, None, None, None) == 1
        assert b"no get_boxes result" in fresh._lib.fmmh_last_error()
    finally:
        fresh.close()
