"""Learned per-pair M2L orders: the GPU training kernels against the CPU
oracle (oracle/mlp.py, pinned by finite differences), the engine's own
truncation-error measurements against a numpy recomputation, and the
selector inside field_fmm against full-order FMM and direct sums.

Bars: FP32 network gradients and outputs within 2e-4 of the float64 oracle
(relative 2-norm); measured truncation errors equal to the numpy ones to
1e-9 relative (FP64 both sides); a selected-order call meets its tolerance
against the undamped direct sum (the far-field semantics of the reference)
within a factor 2, or falls back to expansion_order."""
import math

import numpy as np
import pytest

import oracle
from oracle import mlp
from conftest import rel
from fmm_amd import synth

pytestmark = pytest.mark.gpu


def _water(n_mol, seed, dipole=0.05):
    x, q, excl, df, rng = synth.water_box(n_mol, seed)
    n = x.shape[0]
    m = np.zeros((n, 4))
    m[:, 0] = q
    m[:, 1:4] = rng.normal(0.0, dipole, (n, 3))
    return x, np.asfortranarray(m), excl, df


def _random_case(seed, n):
    rng = np.random.default_rng(seed)
    W = mlp.init(rng, ymean=4.0).astype(np.float32)
    W[256:320] = rng.normal(0.0, 0.3, 64)
    W[4416:4480] = rng.normal(0.0, 0.3, 64)
    norm = np.concatenate([rng.normal(0.0, 1.0, 4), rng.uniform(0.5, 2.0, 4)]).astype(np.float32)
    X = rng.normal(0.0, 1.5, (n, 4)).astype(np.float32)
    Y = rng.uniform(0.0, 11.0, n).astype(np.float32)
    return W, norm, X, Y


@pytest.mark.parametrize("n", [1, 64, 300, 1024])
def test_network_gradient_matches_oracle(gpu_ctx, n):
    W, norm, X, Y = _random_case(n, n)
    loss, g = gpu_ctx.selector_grad(W, norm, X, Y, asym=4.0)
    loss_o, g_o = mlp.loss_grad(W, norm, X, Y, asym=4.0)
    assert abs(loss - loss_o) <= 2e-5 * max(1.0, loss_o)
    assert rel(g, g_o) < 2e-4
    for lo, hi in ((0, 256), (256, 320), (320, 4416), (4416, 4480), (4480, 4544), (4544, 4545)):
        assert rel(g[lo:hi], g_o[lo:hi]) < 1e-3, (lo, hi)


def test_network_forward_matches_oracle(gpu_ctx):
    W, norm, X, _ = _random_case(3, 5000)
    assert rel(gpu_ctx.selector_predict(W, norm, X), mlp.forward(W, norm, X)) < 2e-5


def _moments(xs, ms, c, cap):
    """Cartesian moments M(s) = sum_sites sum_{k <= s} C(s, k) d^{s-k} m(k), |s| <= cap, about c."""
    comps = [(a, b, cc) for k in range(cap + 1) for a in range(k, -1, -1) for b in range(k - a, -1, -1)
             for cc in [k - a - b]]
    dim = ms.shape[1]
    src = comps[:dim]
    d = xs - c
    out = np.zeros(len(comps))
    for si, s in enumerate(comps):
        acc = 0.0
        for ki, k in enumerate(src):
            if k[0] > s[0] or k[1] > s[1] or k[2] > s[2]:
                continue
            coef = math.comb(s[0], k[0]) * math.comb(s[1], k[1]) * math.comb(s[2], k[2])
            mono = d[:, 0] ** (s[0] - k[0]) * d[:, 1] ** (s[1] - k[1]) * d[:, 2] ** (s[2] - k[2])
            acc += coef * np.sum(mono * ms[:, ki])
        out[si] = acc
    return comps, out


def _idx(a, b, c):
    k = a + b + c
    return k * (k + 1) * (k + 2) // 6 + (b * b + 2 * b * c + b + c * c + 3 * c) // 2


def _cs(s):
    return (-1.0) ** (sum(s) + 1) / (math.factorial(s[0]) * math.factorial(s[1]) * math.factorial(s[2]))


def test_measured_truncation_errors_match_numpy(gpu_ctx):
    """The training data: for sampled M2L pairs, the engine's error of the
    pair's contribution at every order q <= cap (M2L truncated at q, exact
    evaluation at up to 32 sites of the target box) against the direct sum
    over the source box, recomputed here with the oracle's tensors."""
    from fmm_amd.octree import Context
    x, m, _, _ = _water(1000, 11)
    cap, theta, ncrit = 6, 0.5, 32
    ctx = Context(0)
    try:
        ctx.train_selector(x, m, theta, ncrit, cap, nsamples=64, epochs=1, seed=3)
        pairs, feat, err, fref = ctx.selector_samples(64, cap)
        t = ctx.tree(x, ncrit, theta)
    finally:
        ctx.close()
    ms = m * np.array([_cs(s) for s in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]])
    for smp in range(0, 64, 8):
        i, j = (int(v) for v in pairs[smp])
        src = t["members"][t["mem_off"][j]:t["mem_off"][j] + t["nleaf"][j]]
        ci = t["nleaf"][i]
        nt = min(32, ci)
        tgt = np.array([t["members"][t["mem_off"][i] + (k * ci) // nt] for k in range(nt)])
        # exact field of the source box at the target sites
        exact = np.zeros((nt, 4))
        for a, ti in enumerate(tgt):
            d = x[ti] - x[src]
            T = oracle.tensor(2, "", d[:, 0], d[:, 1], d[:, 2])
            for c, tc in enumerate([(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]):
                for k, sk in enumerate([(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]):
                    exact[a, c] += np.sum(ms[src, k] * T[:, _idx(*(np.add(sk, tc)))])
        comps, M = _moments(x[src], m[src], t["centers"][j], cap)
        R = t["centers"][i] - t["centers"][j]
        T = oracle.tensor(2 * cap, "", R[:1], R[1:2], R[2:3])[0]
        dd = x[tgt] - t["centers"][i]
        for q in range(cap + 1):
            nq = (q + 1) * (q + 2) * (q + 3) // 6
            L = np.zeros(nq)
            for ti_, tc in enumerate(comps[:nq]):
                L[ti_] = sum(_cs(s) * M[si] * T[_idx(*(np.add(s, tc)))] for si, s in enumerate(comps[:nq]))
            f = np.zeros((nt, 4))
            for e in comps[:nq]:
                mono = (dd[:, 0] ** e[0] * dd[:, 1] ** e[1] * dd[:, 2] ** e[2] /
                        (math.factorial(e[0]) * math.factorial(e[1]) * math.factorial(e[2])))
                for c, tc in enumerate([(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]):
                    tt = tuple(np.add(e, tc))
                    if sum(tt) <= q:
                        f[:, c] += mono * L[_idx(*tt)]
            e_np = np.sqrt(np.mean(np.sum((f[:, 1:] - exact[:, 1:]) ** 2, axis=1)))
            assert abs(err[smp, q] - e_np) <= 1e-9 * e_np + 1e-12 * fref[smp], (smp, q, err[smp, q], e_np)
        assert abs(fref[smp] - np.sqrt(np.mean(np.sum(exact[:, 1:] ** 2, axis=1)))) <= 1e-10 * fref[smp]
        Rn = np.linalg.norm(R)
        assert np.isclose(feat[smp, 0], t["r"][i] / Rn, rtol=1e-6)
        assert np.isclose(feat[smp, 1], t["r"][j] / Rn, rtol=1e-6)


@pytest.fixture(scope="module")
def trained():
    """A selector trained on one water box (cap 10, theta 0.4)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from fmm_amd.octree import Context
    x, m, _, _ = _water(3000, 21)
    ctx = Context(0)
    info = ctx.train_selector(x, m, 0.4, 64, 10, nsamples=4096, epochs=60, seed=5)
    loss = ctx.selector_loss()
    W, norm, meta = ctx.selector_weights()
    yield ctx, info, loss, (W, norm, meta)
    ctx.close()


def test_training_reduces_the_loss(trained):
    ctx, info, loss, _ = trained
    assert info["examples"] >= 4096 * 8
    assert info["graph_replays"] >= 1
    assert np.all(np.isfinite(loss))
    first, last = float(np.mean(loss[:8])), float(np.mean(loss[-16:]))
    print(f"training: {info}, loss {first:.3f} -> {last:.4f}")
    assert last < 0.1 * first
    assert last < 0.5    # mean asymmetric squared error in orders


def test_trained_weights_reproduce_the_oracle(trained):
    """The GPU-trained network, evaluated by the oracle and by the inference
    kernel, gives the same outputs."""
    ctx, _, _, (W, norm, meta) = trained
    rng = np.random.default_rng(0)
    X = np.column_stack([rng.uniform(0.05, 0.4, 2000), rng.uniform(0.05, 0.4, 2000),
                         rng.uniform(-4, 0, 2000), rng.uniform(-12, -3, 2000)]).astype(np.float32)
    assert rel(ctx.selector_predict(W, norm, X), mlp.forward(W, norm, X)) < 2e-5


def _fmm(ctx, x, m, e, theta, p, F=1):
    return ctx.field_fmm(0, x, m, e, theta, 64, p, F)


@pytest.mark.parametrize("tol", [1e-4, 1e-5])
def test_learned_orders_meet_tolerance_and_save_m2l_work(trained, tol):
    """On another water box than the one trained on: the learned orders keep
    the far-field error near the requested tolerance (against the undamped
    direct sum, which is the reference's far-field semantics) and cut the M2L
    work well below expansion_order's -- the gain of the network over the
    method without it."""
    from fmm_amd.octree import Context
    _, _, _, (W, norm, meta) = trained
    x, m, e, _ = _water(5000, 33)
    ctx = Context(0)
    try:
        ctx.set_selector_weights(W, norm, meta)
        direct = ctx.field_direct(0, x, m, e, 1)
        full = _fmm(ctx, x, m, e, 0.4, 10)
        err_full = rel(full[:, 1:4], direct[:, 1:4])
        ctx.set_order_tolerance(tol, spot_checks=256)
        sel = _fmm(ctx, x, m, e, 0.4, 10)
        st = ctx.selection_stats()
        err_sel = rel(sel[:, 1:4], direct[:, 1:4])
        print(f"tol {tol:g}: full-order error {err_full:.3e}, learned {err_sel:.3e}, stats {st}")
        assert st["active"]
        assert st["spot_error"] >= 0.0
        if st["fallback"]:
            assert np.array_equal(sel, full)
        else:
            assert err_sel <= 2.0 * tol + err_full
            assert st["m2l_flops"] < 0.5 * st["m2l_flops_full"]
    finally:
        ctx.close()


def test_impossible_tolerance_falls_back_to_expansion_order(trained):
    from fmm_amd.octree import Context
    _, _, _, (W, norm, meta) = trained
    x, m, e, _ = _water(2000, 41)
    ctx = Context(0)
    try:
        ctx.set_selector_weights(W, norm, meta)
        full = _fmm(ctx, x, m, e, 0.5, 6)
        ctx.set_order_tolerance(1e-13, spot_checks=64)
        sel = _fmm(ctx, x, m, e, 0.5, 6)
        st = ctx.selection_stats()
        assert st["fallback"] and st["spot_error"] > 1e-13
        assert np.array_equal(sel, full)
        # the full-order result misses 1e-13 too: the tolerance is out of reach, nothing is tightened
        assert st["full_order_spot_error"] > 1e-13
        assert st["tolerance_scale"] == 1.0
    finally:
        ctx.close()


def test_fallback_tightens_later_calls_when_full_order_meets_tolerance(trained):
    """Per-pair tolerances 1000x too loose: the learned orders miss 1e-4, the
    full-order result meets it, so later calls get a per-pair tolerance
    tighter by about the factor missed (at most 1/2, at least 1/8 per
    fallback); setting the tolerance again resets it."""
    from fmm_amd.octree import Context
    _, _, _, (W, norm, meta) = trained
    x, m, e, _ = _water(2000, 43)
    ctx = Context(0)
    try:
        ctx.set_selector_weights(W, norm, meta)
        full = _fmm(ctx, x, m, e, 0.4, 8)
        ctx.set_order_tolerance(1e-4, spot_checks=256, safety=1000.0)
        sel = _fmm(ctx, x, m, e, 0.4, 8)
        st = ctx.selection_stats()
        assert st["fallback"] and st["spot_error"] > 1e-4
        assert 0.0 <= st["full_order_spot_error"] <= 1e-4
        assert np.array_equal(sel, full)
        assert 0.125 - 1e-15 <= st["tolerance_scale"] <= 0.5
        _fmm(ctx, x, m, e, 0.4, 8)
        st2 = ctx.selection_stats()
        assert st2["tau"] == pytest.approx(st["tau"] * st["tolerance_scale"], rel=1e-12)
        ctx.set_order_tolerance(1e-4, spot_checks=256, safety=1000.0)
        assert ctx.selection_stats()["tolerance_scale"] == 1.0
    finally:
        ctx.close()


def test_without_a_trained_selector_every_pair_keeps_expansion_order():
    from fmm_amd.octree import Context
    x, m, e, _ = _water(2000, 43)
    ctx = Context(0)
    try:
        full = _fmm(ctx, x, m, e, 0.5, 6)
        ctx.set_order_tolerance(1e-4, spot_checks=16)
        same = _fmm(ctx, x, m, e, 0.5, 6)
        assert not ctx.selection_stats()["active"]
        assert np.array_equal(same, full)
    finally:
        ctx.close()


def test_force_full_orders_match_plain_call(trained):
    """The selection machinery with every pair at expansion_order (the (class,
    order) grouped tiles then equal the class tiles) reproduces the plain call."""
    from fmm_amd.octree import Context
    _, _, _, (W, norm, meta) = trained
    x, m, e, _ = _water(2000, 47)
    ctx = Context(0)
    try:
        ctx.set_selector_weights(W, norm, meta)
        full = _fmm(ctx, x, m, e, 0.5, 8)
        ctx.set_order_tolerance(1e-6, force_full=True)
        same = _fmm(ctx, x, m, e, 0.5, 8)
        st = ctx.selection_stats()
        assert st["active"] and st["order_hist"][8] == st["order_hist"].sum()
        assert rel(same, full) < 1e-14
    finally:
        ctx.close()


def test_learned_orders_with_many_sets(trained):
    """Many multipole sets (per-target MFMA M2L): per-pair orders from the
    largest source norm over the sets; every set meets the tolerance."""
    from fmm_amd.octree import Context
    _, _, _, (W, norm, meta) = trained
    x, m, e, _ = _water(2000, 53)
    n = x.shape[0]
    rng = np.random.default_rng(9)
    S = 16
    M = np.zeros((n, 4, S), order="F")
    for k in range(S):
        M[:, 0, k] = m[:, 0]
        M[:, 1:4, k] = rng.normal(0.0, 0.05, (n, 3))
    ctx = Context(0)
    try:
        ctx.set_selector_weights(W, norm, meta)
        full = ctx.field_fmm_sets(0, x, M, e, 0.4, 64, 8, 1)
        ctx.set_order_tolerance(1e-4)
        sel = ctx.field_fmm_sets(0, x, M, e, 0.4, 64, 8, 1)
        st = ctx.selection_stats()
        assert st["active"] and st["m2l_flops"] < 0.6 * st["m2l_flops_full"]
        for k in (0, S - 1):
            d = ctx.field_direct(0, x, np.asfortranarray(M[:, :, k]), e, 1)
            assert rel(sel[:, 1:4, k], d[:, 1:4]) <= 2e-4 + rel(full[:, 1:4, k], d[:, 1:4])
    finally:
        ctx.close()


@pytest.mark.parametrize("S", [2, 6])
def test_learned_orders_with_unlike_sets(trained, S):
    """Sets of very different size in one call (the last set: dipoles 1000x smaller, no charges): the
    per-pair tolerance comes from the smallest near-field RMS over the sets and the spot checks cover
    every set, so every set meets the tolerance relative to its own field, or the call falls back to
    exactly the full-order result."""
    from fmm_amd.octree import Context
    _, _, _, (W, norm, meta) = trained
    x, m, e, _ = _water(2000, 59)
    n = x.shape[0]
    rng = np.random.default_rng(13)
    M = np.zeros((n, 4, S), order="F")
    for k in range(S - 1):
        M[:, 0, k] = m[:, 0]
        M[:, 1:4, k] = rng.normal(0.0, 0.05, (n, 3))
    M[:, 1:4, S - 1] = rng.normal(0.0, 5e-5, (n, 3))
    ctx = Context(0)
    try:
        ctx.set_selector_weights(W, norm, meta)
        full = ctx.field_fmm_sets(0, x, M, e, 0.4, 64, 8, 1)
        ctx.set_order_tolerance(1e-4, spot_checks=256)
        sel = ctx.field_fmm_sets(0, x, M, e, 0.4, 64, 8, 1)
        st = ctx.selection_stats()
        assert st["active"]
        if st["fallback"]:
            assert np.array_equal(sel, full)
        else:
            assert st["spot_error"] <= 1e-4
            for k in range(S):
                d = ctx.field_direct(0, x, np.asfortranarray(M[:, :, k]), e, 1)
                assert rel(sel[:, 1:4, k], d[:, 1:4]) <= 2e-4 + rel(full[:, 1:4, k], d[:, 1:4]), k
    finally:
        ctx.close()


def test_uneven_batch_of_systems_meets_tolerance():
    """The uneven batch at 1/50 scale: water systems of uneven size, 8 THOLE dipole
    sets each, learned orders (cap 10, tolerance 1e-6, spot checks) in mixed
    precision with a selector trained at the batch's theta.  Every system
    either passes its spot check, with every set within 2e-6 of the same call
    at full order, or falls back to exactly the full-order result."""
    from fmm_amd.octree import Context
    c = synth.config(4, scale=0.02)
    sizes = np.array([sy["coordinates"].shape[0] for sy in c["systems"]])
    order = np.argsort(sizes)
    th, cap = c["theta"], c["expansion_order"]
    xt, mt, _, _ = _water(3000, 61)
    ctx = Context(0)
    try:
        ctx.train_selector(xt, mt, th, c["ncrit"], cap, nsamples=2048, epochs=40, seed=7)
        ctx.set_precision(True)
        for idx in (order[0], order[20], order[44], order[-1]):
            sy = c["systems"][idx]
            x, m, e, df = sy["coordinates"], sy["multipoles"], sy["exclusions"], sy["damping_factors"]
            ctx.set_order_tolerance(0.0)
            full = ctx.field_fmm_sets(0, x, m, e, th, c["ncrit"], cap, 1, "THOLE", df)
            ctx.set_order_tolerance(c["order_tolerance"], spot_checks=256)
            got = ctx.field_fmm_sets(0, x, m, e, th, c["ncrit"], cap, 1, "THOLE", df)
            st = ctx.selection_stats()
            if ctx.stats()["n_m2l"] == 0:
                # too small for any far pair: nothing to select, the plain result
                assert not st["active"] and np.array_equal(got, full)
                continue
            assert st["active"] and st["spot_error"] >= 0.0, (sizes[idx], st)
            if st["fallback"]:
                assert np.array_equal(got, full)
            else:
                assert st["spot_error"] <= c["order_tolerance"]
                for k in range(c["sets"]):
                    assert rel(got[:, 1:4, k], full[:, 1:4, k]) <= 2e-6, (sizes[idx], k)
    finally:
        ctx.close()
