"""HIP graph replay of tree-reuse calls (FMMH_OPT_GRAPH): with the tree cache
on, the first call of a shape runs eagerly, the next one captures its device
work as a graph and launches it, and later calls replay the graph.  Every
replay must give the eager call's field bit for bit, also when the multipole
values change in place between calls (the induced-dipole use) and when other
GPU work runs between the replays; any change of shape, expansion order,
option or buffer size must fall back to an eager call.  Every call goes
through the C ABI of libfmm_hip.so."""
import numpy as np
import pytest

from conftest import load_golden
from fmm_amd import synth

pytestmark = pytest.mark.gpu


def _context(graph, mixed=False):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from fmm_amd.octree import Context
    c = Context(0)
    c.set_tree_reuse(True)
    c.set_precision(mixed)
    c.set_graph(graph)
    return c


def _sets(n, dim, nsets, seed):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(np.stack([synth.random_multipoles(rng, n, dim) for _ in range(nsets)], axis=2))


def _other_gpu_work(other, x, m, e):
    """Eager GPU work between replays: a torch product and a field_direct call on another context."""
    import torch
    a = torch.randn(512, 512, device="cuda:0")
    (a @ a).sum().item()
    other.field_direct(0, x[:300], np.asfortranarray(m[:300]), e[:300], 1)
    torch.cuda.synchronize()


@pytest.mark.parametrize("nsets,mixed", [(1, False), (3, False), (20, False), (20, True)])
def test_replays_equal_eager_calls(nsets, mixed):
    g = load_golden("water999_q.npz")
    x, e, df = g["coordinates"], g["exclusions"], g["thole_df"]
    m = _sets(x.shape[0], 4, nsets, 41)

    def call(c):
        if nsets == 1:
            return c.field_fmm(0, x, np.asfortranarray(m[:, :, 0]), e, 0.5, 64, 6, 2, "THOLE", df)
        return c.field_fmm_sets(0, x, m, e, 0.5, 64, 6, 2, "THOLE", df)

    eager, ctx = _context(False, mixed), _context(True, mixed)
    try:
        call(eager)
        ref = call(eager)                       # the tree-reuse path, eager
        modes = []
        for it in range(6):
            got = call(ctx)
            modes.append(ctx.stats()["graph"])
            np.testing.assert_array_equal(got, ref, err_msg=f"call {it}")
            _other_gpu_work(eager, x, m[:, :, 0], e)
        # first call builds the tree (no graph), second runs eagerly, third captures, then replays
        assert modes == [0, 0, 1, 2, 2, 2], modes
        assert ctx.stats()["graph_replays"] == 3
    finally:
        eager.close()
        ctx.close()


def test_replays_read_multipoles_changed_in_place():
    """Device arrays: the graph reads the multipole array at replay time, so
    new values written into the same memory give the eager field of the new
    values (the induced-dipole iteration)."""
    import torch
    g = load_golden("water999_q.npz")
    x, e, df = g["coordinates"], g["exclusions"], g["thole_df"]
    n = x.shape[0]
    dev = torch.device("cuda:0")
    eager, ctx = _context(False), _context(True)
    try:
        X = torch.as_tensor(np.ascontiguousarray(x.T), device=dev).T
        E = torch.as_tensor(np.ascontiguousarray(e.T), device=dev).T
        D = torch.as_tensor(df, device=dev)
        M = torch.zeros((4, n), dtype=torch.float64, device=dev).T
        out = torch.zeros((10, n), dtype=torch.float64, device=dev).T
        rng = np.random.default_rng(43)
        modes = []
        for it in range(7):
            mk = synth.random_multipoles(rng, n, 4)
            M.copy_(torch.as_tensor(mk, device=dev))
            out.zero_()
            torch.cuda.synchronize()
            ctx.field_fmm_device(X, M, E, 0.5, 64, 7, 2, out, "THOLE", D)
            modes.append(ctx.stats()["graph"])
            got = out.cpu().numpy()
            eager.field_fmm(0, x, np.asfortranarray(mk), e, 0.5, 64, 7, 2, "THOLE", df)
            ref = eager.field_fmm(0, x, np.asfortranarray(mk), e, 0.5, 64, 7, 2, "THOLE", df)
            np.testing.assert_array_equal(got, ref, err_msg=f"call {it}")
        assert modes[:4] == [0, 0, 1, 2] and all(v == 2 for v in modes[3:]), modes
    finally:
        eager.close()
        ctx.close()


def test_changes_fall_back_to_eager_calls():
    """Another expansion order on the same tree rebuilds the M2L classes, an
    option change or a larger call reallocates or changes state: each starts
    the eager -> capture -> replay sequence over, with eager-equal fields."""
    g = load_golden("water999_dipole_orders.npz")
    x, m, e = g["coordinates"], g["multipoles"], g["exclusions"]
    eager, ctx = _context(False), _context(True)
    try:
        refs = {}
        for p in (5, 7):
            eager.field_fmm(0, x, m, e, 0.5, 64, p, 1)
            refs[p] = eager.field_fmm(0, x, m, e, 0.5, 64, p, 1)
        seq = [5, 5, 5, 5, 7, 7, 5, 5, 5, 5]
        modes = []
        for p in seq:
            got = ctx.field_fmm(0, x, m, e, 0.5, 64, p, 1)
            modes.append(ctx.stats()["graph"])
            np.testing.assert_array_equal(got, refs[p], err_msg=f"p={p}")
        assert modes == [0, 0, 1, 2, 0, 0, 0, 0, 1, 2], modes
        ctx.set_timing(True)                    # timing events: eager
        got = ctx.field_fmm(0, x, m, e, 0.5, 64, 5, 1)
        assert ctx.stats()["graph"] == 0
        np.testing.assert_array_equal(got, refs[5])
        ctx.set_timing(False)
        modes = []
        for _ in range(3):
            got = ctx.field_fmm(0, x, m, e, 0.5, 64, 5, 1)
            modes.append(ctx.stats()["graph"])
            np.testing.assert_array_equal(got, refs[5])
        assert modes == [0, 1, 2], modes
    finally:
        eager.close()
        ctx.close()
