"""Tree / list reuse across field_fmm calls (FMMH_OPT_REUSE_TREE): a cached
tree must give bit-for-bit the fields of a rebuild, and any change of
coordinates, ncrit or theta must rebuild.  Every call goes through the C ABI
of libfmm_hip.so."""
import numpy as np
import pytest

from conftest import load_golden
from fmm_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture
def ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from fmm_amd.octree import Context
    c = Context(0)
    c.set_timing(True)
    yield c
    c.close()


def _fresh(x, m, e, theta, ncrit, p, F, damp="", df=()):
    from fmm_amd.octree import Context
    c = Context(0)
    try:
        return c.field_fmm(0, x, m, e, theta, ncrit, p, F, damp, df)
    finally:
        c.close()


def test_reuse_is_bit_identical_and_skips_the_build(ctx):
    g = load_golden("water999_q.npz")
    x, e, df = g["coordinates"], g["exclusions"], g["thole_df"]
    rng = np.random.default_rng(5)
    ctx.set_tree_reuse(True)
    for it in range(3):
        m = np.asfortranarray(synth.random_multipoles(rng, x.shape[0], 4))
        got = ctx.field_fmm(0, x, m, e, 0.5, 64, 6, 2, "THOLE", df)
        phases = ctx.phase_times()
        assert ("tree_reused" in phases) == (it > 0)
        assert ("lists" in phases) == (it == 0)
        np.testing.assert_array_equal(got, _fresh(x, m, e, 0.5, 64, 6, 2, "THOLE", df))
    assert ctx.stats()["p2p_site_pairs"] > 0


def test_changed_inputs_rebuild(ctx):
    g = load_golden("water999_dipole_orders.npz")
    x, m, e = g["coordinates"], g["multipoles"], g["exclusions"]
    ctx.set_tree_reuse(True)
    ctx.field_fmm(0, x, m, e, 0.5, 64, 5, 1)
    # one coordinate moved by one ulp
    x2 = np.array(x, order="F")
    x2[17, 1] = np.nextafter(x2[17, 1], np.inf)
    cases = [(x2, 0.5, 64, 5), (x2, 0.6, 64, 5), (x2, 0.6, 16, 5), (x2, 0.6, 16, 7)]
    for xx, theta, ncrit, p in cases:
        got = ctx.field_fmm(0, xx, m, e, theta, ncrit, p, 1)
        phases = ctx.phase_times()
        rebuilt = "lists" in phases
        # only the expansion order changed in the last case: tree kept, classes recomputed
        assert rebuilt == (p == 5), (theta, ncrit, p)
        np.testing.assert_array_equal(got, _fresh(xx, m, e, theta, ncrit, p, 1))


def test_tree_export_and_option_off_drop_the_cache(ctx):
    g = load_golden("water999_dipole_orders.npz")
    x, m, e = g["coordinates"], g["multipoles"], g["exclusions"]
    ctx.set_tree_reuse(True)
    ctx.field_fmm(0, x, m, e, 0.5, 64, 5, 1)
    ctx.tree(x[:500], 16, 0.5)          # overwrites the tree buffers
    got = ctx.field_fmm(0, x, m, e, 0.5, 64, 5, 1)
    assert "lists" in ctx.phase_times()
    np.testing.assert_array_equal(got, _fresh(x, m, e, 0.5, 64, 5, 1))
    ctx.set_tree_reuse(False)
    ctx.field_fmm(0, x, m, e, 0.5, 64, 5, 1)
    assert "lists" in ctx.phase_times()


def test_reuse_with_sets_and_device_calls(ctx):
    import torch
    g = load_golden("water999_q.npz")
    x, e, df = g["coordinates"], g["exclusions"], g["thole_df"]
    rng = np.random.default_rng(9)
    m = np.asfortranarray(np.stack([synth.random_multipoles(rng, x.shape[0], 10) for _ in range(18)], 2))
    ctx.set_tree_reuse(True)
    dev = torch.device("cuda:0")
    X = torch.as_tensor(np.ascontiguousarray(x.T), device=dev).T
    E = torch.as_tensor(np.ascontiguousarray(e.T), device=dev).T
    D = torch.as_tensor(df, device=dev)
    M = torch.as_tensor(np.ascontiguousarray(m.transpose(2, 1, 0)), device=dev).permute(2, 1, 0)
    outs = []
    for _ in range(2):
        out = torch.zeros((18, 4, x.shape[0]), dtype=torch.float64, device=dev).permute(2, 1, 0)
        ctx.field_fmm_sets_device(X, M, E, 0.5, 64, 6, 1, out, "AMOEBA", D)
        outs.append(out.cpu().numpy())
    assert "tree_reused" in ctx.phase_times()
    np.testing.assert_array_equal(outs[0], outs[1])


def test_host_waits_per_call(ctx):
    """Host waits on the device inside one field_fmm call (stats host_waits): a call that builds its tree
    waits five times -- the node count (to size the node arrays), the tree's depth and level sizes, the
    interaction-list sizes, the near-field workgroup order and M2L class counts (one round trip), and the
    call's completion; a call that reuses the tree waits for the coordinate comparison and its completion."""
    g = load_golden("water999_q.npz")
    x, e, df = g["coordinates"], g["exclusions"], g["thole_df"]
    m = np.asfortranarray(synth.random_multipoles(np.random.default_rng(9), x.shape[0], 4))
    ctx.set_tree_reuse(True)
    ctx.field_fmm(0, x, m, e, 0.5, 64, 6, 2, "THOLE", df)
    built = ctx.stats()["host_waits"]
    ctx.field_fmm(0, x, m, e, 0.5, 64, 6, 2, "THOLE", df)
    reused = ctx.stats()["host_waits"]
    print(f"host waits: tree built {built}, tree reused {reused}")
    assert built == 5
    assert reused == 2
