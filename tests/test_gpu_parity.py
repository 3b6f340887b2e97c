"""HIP engine vs the reference's golden vectors and the oracle.  Every call
goes through the C ABI of libfmm_hip.so (fmm_amd.octree -> ctypes).

Bars: <= 1e-12 relative 2-norm error per output-order block for FP64 fields;
exact equality for the tree and interaction lists (integer work and the FP64
node-centre chain)."""
import numpy as np
import pytest

import oracle
from conftest import block_rel, load_golden, rel
from fmm_amd import synth

pytestmark = pytest.mark.gpu
PARITY = 1e-12


def test_uniform4096_direct(gpu_ctx):
    g = load_golden("uniform4096_p5.npz")
    fd = gpu_ctx.field_direct(0, g["coordinates"], g["multipoles"], g["exclusions"], 1)
    assert max(block_rel(fd, g["field_direct"], 1)) < PARITY


def test_uniform4096_fmm(gpu_ctx):
    g = load_golden("uniform4096_p5.npz")
    ff = gpu_ctx.field_fmm(0, g["coordinates"], g["multipoles"], g["exclusions"], 0.5, 64, 5, 1)
    assert max(block_rel(ff, g["field_fmm"], 1)) < PARITY


@pytest.mark.parametrize("damp", ["", "THOLE", "AMOEBA", "ERF"])
def test_water_direct_damped(gpu_ctx, damp):
    g = load_golden("water999_q.npz")
    df = g["erf_df"] if damp == "ERF" else g["thole_df"]
    key = damp.lower() or "none"
    fd = gpu_ctx.field_direct(0, g["coordinates"], g["multipoles"], g["exclusions"], 2, damp, df)
    assert max(block_rel(fd, g[f"direct_{key}"], 2, damp == "AMOEBA")) < PARITY


@pytest.mark.parametrize("damp", ["", "THOLE", "AMOEBA", "ERF"])
def test_water_fmm_damped(gpu_ctx, damp):
    g = load_golden("water999_q.npz")
    df = g["erf_df"] if damp == "ERF" else g["thole_df"]
    key = damp.lower() or "none"
    ff = gpu_ctx.field_fmm(0, g["coordinates"], g["multipoles"], g["exclusions"], 0.5, 64, 6, 2, damp, df)
    assert max(block_rel(ff, g[f"fmm_{key}"], 2)) < PARITY


@pytest.mark.parametrize("F", [0, 1])
def test_lower_field_orders_are_leading_blocks(gpu_ctx, F):
    g = load_golden("water999_q.npz")
    nf = synth.ncomp(F)
    ff = gpu_ctx.field_fmm(0, g["coordinates"], g["multipoles"], g["exclusions"], 0.5, 64, 6, F, "THOLE",
                           g["thole_df"])
    assert max(block_rel(ff, g["fmm_thole"][:, :nf], F)) < PARITY


@pytest.mark.parametrize("p", [2, 3, 4, 5, 6])
def test_water_dipole_orders(gpu_ctx, p):
    g = load_golden("water999_dipole_orders.npz")
    ff = gpu_ctx.field_fmm(0, g["coordinates"], g["multipoles"], g["exclusions"], 0.5, 64, p, 1)
    assert max(block_rel(ff, g[f"fmm_p{p}"], 1)) < PARITY


@pytest.mark.parametrize("p", [7, 8, 10])
def test_high_orders_vs_oracle(gpu_ctx, p):
    """Above p = 6 the reference's own T tensors overflow; the oracle (exact
    tensors) is the bar there."""
    g = load_golden("water999_dipole_orders.npz")
    x, m, e = g["coordinates"], g["multipoles"], g["exclusions"]
    ff = gpu_ctx.field_fmm(0, x, m, e, 0.5, 64, p, 1)
    assert max(block_rel(ff, oracle.field_fmm(x, m, e, 0.5, 64, p, 1), 1)) < PARITY


def test_theta06_ncrit16(gpu_ctx):
    g = load_golden("water999_dipole_orders.npz")
    ff = gpu_ctx.field_fmm(0, g["coordinates"], g["multipoles"], g["exclusions"], 0.6, 16, 5, 1)
    assert max(block_rel(ff, g["fmm_theta06_ncrit16_p5"], 1)) < PARITY


@pytest.mark.parametrize("name", ["charges_tiny40.npz", "charges_small300.npz"])
def test_charges(gpu_ctx, name):
    g = load_golden(name)
    x, q, e = g["coordinates"], g["multipoles"], g["exclusions"]
    assert max(block_rel(gpu_ctx.field_direct(0, x, q, e, 2), g["field_direct"], 2)) < PARITY
    assert max(block_rel(gpu_ctx.field_fmm(0, x, q, e, 0.5, 64, 4, 2), g["field_fmm"], 2)) < PARITY


def _tree_equal(a, b):
    for k in ("depth", "parent", "is_leaf", "nleaf", "mem_off", "members", "m2l_off", "m2l_src", "p2p_off",
              "p2p_src"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    for k in ("centers", "r", "rmax"):
        np.testing.assert_array_equal(a[k].view(np.uint64), b[k].view(np.uint64), err_msg=k)
    assert a["max_depth"] == b["max_depth"]


@pytest.mark.parametrize("case", ["uniform", "water", "clusters", "ties"])
def test_tree_and_lists_bit_exact(gpu_ctx, case):
    if case == "uniform":
        x = synth.config(0)["coordinates"]
        ncrit, theta = 64, 0.5
    elif case == "water":
        x = synth.water_box(3000, 3)[0]
        ncrit, theta = 64, 0.5
    elif case == "clusters":
        x, _ = synth.gaussian_clusters(16, 512, 100.0, 3.0, 9)
        ncrit, theta = 64, 0.6
    else:   # lattice points: every theta = 0.5 MAC tie of the lattice is hit
        g = np.stack(np.meshgrid(*[np.arange(24.0)] * 3, indexing="ij"), -1).reshape(-1, 3)
        x = np.asfortranarray(g * 1.5 + 0.25)
        ncrit, theta = 16, 0.5
    _tree_equal(gpu_ctx.tree(x, ncrit, theta), oracle.tree(x, ncrit, theta))


def test_list_overflow_retry_gives_same_lists_and_field(gpu_ctx):
    """Tiny initial list buffers: every traversal step overflows, the kernels
    read and write only inside the buffers, and the host regrows them."""
    from fmm_amd.octree import Context
    g = load_golden("uniform4096_p5.npz")
    ctx = Context(0)
    try:
        ctx.set_list_capacity(1e-4)
        ff = ctx.field_fmm(0, g["coordinates"], g["multipoles"], g["exclusions"], 0.5, 64, 5, 1)
        assert ctx.stats()["list_retries"] >= 2
        assert max(block_rel(ff, g["field_fmm"], 1)) < PARITY
        _tree_equal(ctx.tree(g["coordinates"], 64, 0.5), oracle.tree(g["coordinates"], 64, 0.5))
    finally:
        ctx.close()


def test_field_is_accumulated(gpu_ctx):
    g = load_golden("charges_small300.npz")
    x, q, e = g["coordinates"], g["multipoles"], g["exclusions"]
    base = np.asfortranarray(np.random.default_rng(0).normal(size=(x.shape[0], 10)))
    f = gpu_ctx.field_fmm(0, x, q, e, 0.5, 64, 4, 2, field=base.copy(order="F"))
    assert rel(f - base, g["field_fmm"]) < PARITY


def test_device_entry_points_match_host(gpu_ctx):
    import torch
    g = load_golden("water999_q.npz")
    dev = torch.device("cuda:0")

    def col(a, dt=torch.float64):
        return torch.as_tensor(np.ascontiguousarray(np.asarray(a).T), dtype=dt, device=dev).T

    X, M = col(g["coordinates"]), col(g["multipoles"])
    E = col(g["exclusions"], torch.int32)
    df = torch.as_tensor(g["thole_df"], device=dev)
    out = col(np.zeros((X.shape[0], 10)))
    gpu_ctx.field_fmm_device(X, M, E, 0.5, 64, 6, 2, out, "THOLE", df)
    ref = gpu_ctx.field_fmm(0, g["coordinates"], g["multipoles"], g["exclusions"], 0.5, 64, 6, 2, "THOLE",
                            g["thole_df"])
    np.testing.assert_array_equal(out.cpu().numpy(), ref)
    out.zero_()
    gpu_ctx.field_direct_device(X, M, E, 2, out, "THOLE", df)
    assert max(block_rel(out.cpu().numpy(), g["direct_thole"], 2)) < PARITY


def test_errors(gpu_ctx):
    from fmm_amd.octree import FmmError
    g = load_golden("charges_small300.npz")
    x, q, e = g["coordinates"], g["multipoles"], g["exclusions"]
    with pytest.raises(FmmError, match="Unknown damp type: thole"):
        gpu_ctx.field_fmm(0, x, q, e, 0.5, 64, 4, 1, "thole", np.ones(len(x)))
    with pytest.raises(FmmError, match="Not enough coordinates for periodic FMM"):
        gpu_ctx.field_fmm(0, x, q, e, 0.5, 400, 4, 1, is_periodic=True, box_matrix=np.eye(3) * 10)
    bad = e.copy()
    bad[3, 0] = len(x) + 5
    with pytest.raises(FmmError, match="outside"):
        gpu_ctx.field_direct(0, x, q, bad, 1)


def test_missing_self_exclusion_gives_nonfinite_like_reference(gpu_ctx):
    g = load_golden("charges_tiny40.npz")
    x, q = g["coordinates"], g["multipoles"]
    e = np.zeros((len(x), 1), dtype=np.int32, order="F")
    f = gpu_ctx.field_direct(0, x, q, e, 0)
    assert not np.isfinite(f).all()


def test_direct_targets_match_full_direct(gpu_ctx):
    """Spot-check direct sum at a subset of sites equals those rows of the
    full direct sum (last block of 13 targets runs with lane groups, so the
    bar is rounding, not identity)."""
    import torch
    g = load_golden("water999_q.npz")
    dev = torch.device("cuda:0")

    def col(a, dt=torch.float64):
        return torch.as_tensor(np.ascontiguousarray(np.asarray(a).T), dtype=dt, device=dev).T

    x, m, e, df = g["coordinates"], g["multipoles"], g["exclusions"], g["thole_df"]
    tg = np.random.default_rng(4).choice(x.shape[0], 77, replace=False).astype(np.int32)
    out = col(np.zeros((77, 10)))
    gpu_ctx.field_direct_targets_device(col(x), col(m), col(e, torch.int32), torch.as_tensor(tg, device=dev), 2,
                                        out, "THOLE", torch.as_tensor(df, device=dev))
    full = gpu_ctx.field_direct(0, x, m, e, 2, "THOLE", df)
    got = out.cpu().numpy()
    assert max(block_rel(got, full[tg], 2)) < 1e-14
    # the first 64 targets form one full block here; sites that also sit in a full
    # 64-site block of the all-sites sum get the same per-lane arithmetic, hence identity
    same = np.nonzero(tg[:64] < (x.shape[0] // 64) * 64)[0]
    assert same.size >= 50
    np.testing.assert_array_equal(got[same], full[tg[same]])
    from fmm_amd.octree import FmmError
    with pytest.raises(FmmError, match="target index"):
        bad = torch.as_tensor(np.array([0, x.shape[0]], np.int32), device=dev)
        gpu_ctx.field_direct_targets_device(col(x), col(m), col(e, torch.int32), bad, 2, col(np.zeros((2, 10))))


def _adversarial_axis(kind, n, rng):
    if kind == "uniform":
        return rng.uniform(0.0, 300.0, n)
    if kind == "centred":        # running sum keeps crossing binades and zero
        return rng.uniform(-50.0, 50.0, n)
    if kind == "quarters":       # x / ulp(s) lands on half-integers: ties broken by the parity of s
        return np.floor(rng.uniform(0.0, 600.0, n)) / 2.0 + 0.25
    if kind == "wide":           # magnitudes over 30 decades
        return rng.normal(0.0, 1.0, n) * 10.0 ** rng.integers(-15, 15, n)
    if kind == "cancel":         # partial sums far above the final sum
        return np.where(rng.integers(0, 2, n) == 1, 1e16, -1e16 + 1.0)
    return rng.uniform(0.0, 300.0, n) + 1e9   # offset: small addends on a large running sum


@pytest.mark.parametrize("kind", ["uniform", "centred", "quarters", "wide", "cancel", "offset"])
@pytest.mark.parametrize("n", [2, 17, 16385, 200001, 1048576])
def test_root_centre_is_left_to_right_sum(gpu_ctx, kind, n):
    """The root centre is sum(coordinates, 1) / N summed left to right, as
    flang evaluates it (src/fmm.F90:184), bit for bit, whatever the data do
    to the running sum (binade changes, ties, cancellation).  numpy's cumsum
    is that sequential sum.  Degenerate inputs may not make a valid octree
    (FMMH_ERR_TREE); the root centre is computed first either way and read
    from the context's "t_root" buffer."""
    rng = np.random.default_rng(n + len(kind))
    x = np.asfortranarray(np.stack([_adversarial_axis(kind, n, rng) for _ in range(3)], 1))
    try:
        gpu_ctx.tree(x, min(64, n), 0.5)
    except RuntimeError:
        pass
    got = gpu_ctx.debug_buffer("t_root", 3)
    want = np.cumsum(x, axis=0)[-1] / n
    assert np.array_equal(got.view(np.int64), want.view(np.int64)), (got, want)
