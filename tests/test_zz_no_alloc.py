"""fmmh_field_direct_dev_noalloc: the direct sum with its working arrays in the
library's fixed per-device memory (no allocation), against
fmmh_field_direct_dev on the same inputs (bit for bit: the same kernels on the
same data) and against the reference's fixtures (<= 1e-12 per block), on
consecutive calls of different sizes; sizes beyond the fixed memory and a
missing context are refused.  Every call goes through the C ABI."""
import numpy as np
import pytest

from conftest import block_rel, load_golden

pytestmark = pytest.mark.gpu
PARITY = 1e-12


def _dev(a):
    """Column-major device copy of a host array."""
    import torch
    a = np.asarray(a)
    if a.ndim == 1:
        return torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return torch.from_numpy(np.ascontiguousarray(a.T)).cuda().T


def _zeros(n, nf):
    import torch
    return torch.zeros((nf, n), dtype=torch.float64, device="cuda").T


def _both(ctx, x, m, e, F, damp="", df=None):
    X, M, E = _dev(x), _dev(m), _dev(np.asfortranarray(np.asarray(e).reshape(x.shape[0], -1)))
    D = _dev(df) if damp else None
    nf = (F + 1) * (F + 2) * (F + 3) // 6
    ref = ctx.field_direct_device(X, M, E, F, _zeros(x.shape[0], nf), damp, D)
    got = ctx.field_direct_device_noalloc(X, M, E, F, _zeros(x.shape[0], nf), damp, D)
    return got.cpu().numpy(), ref.cpu().numpy()


def test_noalloc_matches_the_allocating_call_on_consecutive_calls(gpu_ctx):
    g = load_golden("water999_q.npz")
    got, ref = _both(gpu_ctx, g["coordinates"], g["multipoles"], g["exclusions"], 2, "THOLE", g["thole_df"])
    np.testing.assert_array_equal(got, ref)
    assert max(block_rel(got, g["direct_thole"], 2)) < PARITY
    # second call, smaller: the fixed memory keeps the first call's contents beyond the new sizes
    t = load_golden("charges_tiny40.npz")
    got, ref = _both(gpu_ctx, t["coordinates"], t["multipoles"], t["exclusions"], 2)
    np.testing.assert_array_equal(got, ref)
    assert max(block_rel(got, t["field_direct"], 2)) < PARITY
    # and a source order above 2 (octupoles, field order 3)
    w = load_golden("water_wide_orders.npz")
    key = "a_d20_f3_erf_n64_p5"
    got, ref = _both(gpu_ctx, w["a_coordinates"], np.asfortranarray(w["a_multipoles"][:, :20]), w["a_exclusions"],
                     3, "ERF", w["a_erf_df"])
    np.testing.assert_array_equal(got, ref)
    assert max(block_rel(got[w["rows_a"]], w[f"direct_{key}"], 3)) < PARITY


def test_noalloc_accumulates_into_field(gpu_ctx):
    g = load_golden("charges_small300.npz")
    x, q, e = g["coordinates"], g["multipoles"], g["exclusions"]
    base = np.asfortranarray(np.random.default_rng(0).normal(size=(x.shape[0], 10)))
    f = _dev(base)
    gpu_ctx.field_direct_device_noalloc(_dev(x), _dev(q), _dev(e), 2, f)
    assert max(block_rel(f.cpu().numpy() - base, g["field_direct"], 2)) < PARITY


def test_noalloc_refuses_sizes_beyond_the_fixed_memory(gpu_ctx):
    import torch
    from fmm_amd.octree import FmmError
    n = 300_000                       # 10 components and 3 exclusions: ~38 MB of working arrays
    X = torch.zeros((3, n), dtype=torch.float64, device="cuda").T
    M = torch.zeros((10, n), dtype=torch.float64, device="cuda").T
    E = torch.zeros((3, n), dtype=torch.int32, device="cuda").T
    with pytest.raises(FmmError, match="exceed the fixed"):
        gpu_ctx.field_direct_device_noalloc(X, M, E, 1, _zeros(n, 4))


def test_noalloc_needs_a_context():
    from fmm_amd import _lib
    lib = _lib.load()
    rc = lib.fmmh_field_direct_dev_noalloc(None, 0, None, 1, None, 1, 1, 1, None, 1, 0, 0, None, 1, 0, None, 0)
    assert rc == 1 and b"needs a context" in lib.fmmh_last_error()
