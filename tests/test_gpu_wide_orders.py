"""Source multipole orders above quadrupole and field orders above 2 on the
GPU (src/fmm.F90:548-621: the reference infers the source order from dim and
builds tensors of order m + F), sites at one position, and the range check of
exclusion entries on the device-array entry points.  Every call goes through
the C ABI of libfmm_hip.so.

Bars: <= 1e-12 relative 2-norm error per output-order block against the
compiled reference's fixtures (tests/golden/make_golden.py: wide_orders,
coincident); AMOEBA direct potentials are skipped (its order-0 tensor is
zero, src/tensors_damped_amoeba.F90:77)."""
import numpy as np
import pytest

import oracle
from conftest import block_rel, load_golden

pytestmark = pytest.mark.gpu
PARITY = 1e-12


def _cases():
    return [str(k) for k in load_golden("water_wide_orders.npz")["cases"]]


def _inputs(g, key):
    box, d, f, damp, n, p = key.split("_")
    dim, F, ncrit, p = int(d[1:]), int(f[1:]), int(n[1:]), int(p[1:])
    damp = "" if damp == "none" else damp.upper()
    df = g[f"{box}_erf_df"] if damp == "ERF" else g[f"{box}_thole_df"]
    m = np.asfortranarray(g[f"{box}_multipoles"][:, :dim])
    return g[f"{box}_coordinates"], m, g[f"{box}_exclusions"], df, F, ncrit, p, damp, g[f"rows_{box}"]


@pytest.mark.parametrize("key", _cases())
def test_wide_direct(gpu_ctx, key):
    g = load_golden("water_wide_orders.npz")
    x, m, e, df, F, ncrit, p, damp, rows = _inputs(g, key)
    print(f"field_direct {key}", flush=True)
    fd = gpu_ctx.field_direct(0, x, m, e, F, damp, df)[rows]
    assert max(block_rel(fd, g[f"direct_{key}"], F, damp == "AMOEBA")) < PARITY


@pytest.mark.parametrize("key", _cases())
def test_wide_fmm(gpu_ctx, key):
    g = load_golden("water_wide_orders.npz")
    x, m, e, df, F, ncrit, p, damp, rows = _inputs(g, key)
    print(f"field_fmm {key}", flush=True)
    ff = gpu_ctx.field_fmm(0, x, m, e, 0.5, ncrit, p, F, damp, df)[rows]
    assert max(block_rel(ff, g[f"fmm_{key}"], F)) < PARITY


@pytest.mark.parametrize("key", ["a_d20_f3_thole_n64_p5", "b_d56_f2_none_n16_p6"])
def test_wide_many_sets_match_single_set_calls(gpu_ctx, key):
    """Three sets of one call against three single-set calls (same kernels, one set per workgroup row)."""
    g = load_golden("water_wide_orders.npz")
    x, m, e, df, F, ncrit, p, damp, _ = _inputs(g, key)
    rng = np.random.default_rng(3)
    sets = [m, np.asfortranarray(-0.5 * m), np.asfortranarray(m * rng.uniform(0.5, 1.5, m.shape))]
    M = np.asfortranarray(np.stack(sets, axis=2))
    many = gpu_ctx.field_fmm_sets(0, x, M, e, 0.5, ncrit, p, F, damp, df)
    for s, ms in enumerate(sets):
        one = gpu_ctx.field_fmm(0, x, ms, e, 0.5, ncrit, p, F, damp, df)
        assert max(block_rel(many[:, :, s], one, F)) < 1e-14


def test_orders_outside_the_compiled_range(gpu_ctx):
    from fmm_amd.octree import FmmError
    g = load_golden("water_wide_orders.npz")
    x, e = g["b_coordinates"], g["b_exclusions"]
    m = np.asfortranarray(g["b_multipoles"][:, :56])   # m = 5
    with pytest.raises(FmmError, match="outside the compiled range"):
        gpu_ctx.field_direct(0, x, m, e, 3)            # m + F = 8


@pytest.mark.parametrize("damp", ["", "THOLE"])
def test_coincident_sites(gpu_ctx, damp):
    """Sites 10/11 share a position and exclude each other: finite, as the reference; sites 20/21 share
    one without excluding each other: non-finite rows, as the reference."""
    g = load_golden("coincident300.npz")
    x, m, e, df = g["coordinates"], g["multipoles"], g["exclusions"], g["thole_df"]
    key = damp.lower() or "none"
    fin = np.ones(x.shape[0], bool)
    fin[[20, 21]] = False
    for got, ref in ((gpu_ctx.field_direct(0, x, m, e, 2, damp, df), g[f"direct_{key}"]),
                     (gpu_ctx.field_fmm(0, x, m, e, 0.5, 16, 5, 2, damp, df), g[f"fmm_{key}"])):
        assert np.isfinite(got[fin]).all()
        assert not np.isfinite(got[~fin]).all(axis=1).any()
        assert max(block_rel(got[fin], ref[fin], 2)) < PARITY


def test_device_exclusion_out_of_range(gpu_ctx):
    """The device-array entry points cannot check exclusion entries on the host: an entry above n is
    reported by the exclusion pass (FMMH_ERR_ARG) instead of read out of bounds."""
    import torch
    from fmm_amd.octree import FmmError
    g = load_golden("charges_small300.npz")
    x, q, e = g["coordinates"], g["multipoles"], g["exclusions"]
    n = x.shape[0]
    bad = np.asfortranarray(np.concatenate([e.reshape(n, -1), np.zeros((n, 1), np.int32)], axis=1))
    bad[7, -1] = n + 5
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a.T)).cuda().T   # column-major device copy
    X, Q, E = dev(x), dev(q), dev(bad)
    f = torch.zeros((1, n), dtype=torch.float64, device="cuda").T
    with pytest.raises(FmmError, match="exclusion index"):
        gpu_ctx.field_direct_device(X, Q, E, 0, f)
    with pytest.raises(FmmError, match="exclusion index"):
        gpu_ctx.field_fmm_device(X, Q, E, 0.5, 16, 4, 0, f)
    # the same call with valid entries still works afterwards
    E = dev(np.asfortranarray(e.reshape(n, -1)))
    f.zero_()
    gpu_ctx.field_fmm_device(X, Q, E, 0.5, 16, 4, 0, f)
    ref = oracle.field_fmm(x, q, e, 0.5, 16, 4, 0)
    assert max(block_rel(f.cpu().numpy(), ref[:, :1], 0)) < PARITY


def test_qmmm_octupole_multipoles(gpu_ctx):
    """get_multipoles with octupole sources (dim 20) against the restatement."""
    g = load_golden("water_wide_orders.npz")
    x = g["a_coordinates"]
    m = np.asfortranarray(np.stack([g["a_multipoles"][:, :20], -g["a_multipoles"][:, :20]], axis=2))
    tg = np.asfortranarray(x.mean(axis=0) + np.array([[0.0, 0.0, 0.0], [2.0, -1.0, 0.5]]))
    near, bc, bm = gpu_ctx.get_multipoles(0, x, m, tg, 0.5, 32, 5)
    rn, rc, rm = oracle.get_multipoles(x, m, tg, 0.5, 32, 5, impl="oracle")
    np.testing.assert_array_equal(near, rn)
    np.testing.assert_array_equal(bc, rc)
    for d in range(bm.shape[2]):
        num = np.linalg.norm(bm[:, :, d] - rm[:, :, d], axis=1)
        assert np.all(num <= 1e-12 * np.linalg.norm(rm[:, :, d], axis=1))
