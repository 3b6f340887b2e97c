"""Benchmark of the field engine on one MI355X.

Workload (default): the plain FP64 FMM on a water-like box -- 66,667 rigid
3-site waters (200,001 sites) on a jittered grid at liquid density, charges
plus random dipoles (dim 4), self + 2 intramolecular exclusions per site,
THOLE damping, theta 0.5, ncrit 64, expansion order 6, field order 2.
Inputs are synthetic (seeded) and resident in HBM before the timed region.

A step is one field_fmm call on the device-resident inputs (tree, lists,
P2M, M2M, P2P, M2L, L2L, L2P).  Prints one JSON line with throughput in
site-field evaluations per second, the roofline of the dominant kernel (P2P)
and the reference CPU path timed on this host on a bounded sample.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--no-cpu] [--no-error] [--config 1]

--config 2 runs the batched form instead: the same water geometry with
AMOEBA damping and 32 quadrupole multipole sets on one tree (one
field_fmm_sets call per step; the M2L runs on the FP64 matrix cores);
its value counts site-field evaluations of every set.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 vector = FP64 matrix peak
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the near-field kernel on the default workload, from
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (one run each) scaled by the same
# counters on a plain 1 GiB device copy (factors in the profile).  The profiler
# cannot run inside this process, so the figure is the recorded measurement of
# profiles/k_near_water200k_thole_p6.md, valid for the near-field sources it
# was measured on: K_NEAR_TRAFFIC_SRC is the sha256 prefix of those sources
# (near_src_sha()); with other sources the bench line reports null.
K_NEAR_TRAFFIC_BYTES = 161808384     # 76.9 MB read + 84.9 MB written
K_NEAR_TRAFFIC_SRC = "fcf67a6ee15079f2"
NEAR_SOURCES = ("fmm_amd/csrc/near_kernels.h", "fmm_amd/csrc/near.hip", "fmm_amd/csrc/tensors.h")


def near_src_sha():
    import hashlib
    h = hashlib.sha256()
    for f in NEAR_SOURCES:
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def c_T(K):
    """Algorithmic FLOPs of the order-K tensor recursion (FMA = 2)."""
    return 9 + 2 * (K + 1) + 3 * sum((k + 1) * (k + 2) // 2 * (K - k + 1) for k in range(1, K + 1))


def p2p_flops_per_pair(m, F, damp, sets=1):
    K = m + F
    ncomp = lambda k: (k + 1) * (k + 2) * (k + 3) // 6  # noqa: E731
    d = 0 if damp == "" else (2 + 4 * (K + 1) if damp == "ERF" else 1 + 4 * (K + 1))
    return c_T(K) + d + sets * 2 * ncomp(m) * ncomp(F)


def col(a, dev, dtype=None):
    import torch
    t = torch.as_tensor(np.ascontiguousarray(np.asarray(a).T), device=dev)
    if dtype is not None:
        t = t.to(dtype)
    return t.T


def col3(a, dev):
    """Column-major N x c x S device tensor."""
    import torch
    return torch.as_tensor(np.ascontiguousarray(np.asarray(a).transpose(2, 1, 0)), device=dev).permute(2, 1, 0)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _host_cpu():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline():
    """The reference Fortran (compiled from its own sources into oracle/_ref) on one host core: value = its
    field_fmm with the shipped flags (-O1 -g) on a 1/10-scale sample of the benchmarked workload; beside it
    the same sample at -O3, and the 4,096-site quadrupole workload in full (field_fmm and field_direct)."""
    import oracle
    from fmm_amd import synth
    if not oracle.have_ref():
        return None
    c = synth.config(1, scale=0.1)
    n = c["coordinates"].shape[0]
    args = (c["coordinates"], c["multipoles"], c["exclusions"], c["theta"], c["ncrit"], c["expansion_order"],
            c["field_order"], c["damp_type"], c["damping_factors"])

    def timed(fn):
        t0 = time.perf_counter()
        fn()
        return time.perf_counter() - t0
    dt = timed(lambda: oracle.field_fmm(*args, impl="ref"))
    extra = {}
    if oracle.have_ref("_o3"):
        dt3 = timed(lambda: oracle.field_fmm(*args, impl="ref_o3"))
        extra["same_sample_O3"] = {"seconds": dt3, "value": n / dt3}
    c0 = synth.config(0)
    n0 = c0["coordinates"].shape[0]
    a0 = (c0["coordinates"], c0["multipoles"], c0["exclusions"])
    t_f = timed(lambda: oracle.field_fmm(*a0, c0["theta"], c0["ncrit"], c0["expansion_order"], c0["field_order"],
                                         impl="ref"))
    t_d = timed(lambda: oracle.field_direct(*a0, c0["field_order"], impl="ref"))
    extra["uniform4096_quadrupole_field_fmm"] = {"seconds": t_f, "value": n0 / t_f}
    extra["uniform4096_quadrupole_field_direct"] = {"seconds": t_d, "value": n0 / t_d}
    model, cores = _host_cpu()
    return {"value": n / dt, "unit": "site-field evals/s", "cores": 1, "kind": "reference",
            "sample": f"reference field_fmm (flang -O1 -g as shipped, serial) on the benchmarked recipe at 1/10 "
                      f"scale: {n} sites, THOLE, theta 0.5, ncrit 64, p 6, field order 2; {dt:.1f} s",
            "host_cpu": model, "host_cores": cores, "more": extra}


def plain_sweep(dev, steps, thetas=(0.25, 0.3, 0.35, 0.4), orders=(7, 8, 9, 10)):
    """The plain FP64 FMM (one set, every pair at expansion_order) on the undamped 200,001-site water
    geometry of target_comparison, over theta x expansion order: ms per call and the relative RMS error of
    the field block against the FP64 direct sum at 4096 seeded sites.  The cheapest setting within 1e-6 is
    the plain side of the comparison."""
    import torch
    from fmm_amd import synth
    from fmm_amd.octree import Context
    c = synth.config(1)
    x = c["coordinates"]
    n = x.shape[0]
    rng = np.random.default_rng(2024)
    m = np.zeros((n, 4), order="F")
    m[:, 0] = c["multipoles"][:, 0]
    m[:, 1:4] = rng.normal(0.0, 0.05, (n, 3))
    X, M, E = col(x, dev), col(m, dev), col(c["exclusions"], dev, torch.int32)
    tg = np.sort(np.random.default_rng(4096).choice(n, 4096, replace=False)).astype(np.int32)
    ctx = Context(0)
    d = col(np.zeros((4096, 4)), dev)
    ctx.field_direct_targets_device(X, M, E, torch.as_tensor(tg, device=dev), 1, d, "", None)
    ref = d.cpu().numpy()
    o = col(np.zeros((n, 4)), dev)
    rows = []
    for th in thetas:
        for p in orders:
            def f():
                o.zero_()
                ctx.field_fmm_device(X, M, E, th, 64, p, 1, o, "", None)
            f()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                f()
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / steps
            err = rel(o.cpu().numpy()[tg, 1:4], ref[:, 1:4])
            rows.append({"theta": th, "expansion_order": p, "ms_per_call": ms, "error": err})
            print(f"plain sweep theta {th} p {p}: {ms:.2f} ms, error {err:.3e}", file=sys.stderr, flush=True)
    ctx.close()
    ok = [r for r in rows if r["error"] <= 1e-6]
    best = min(ok, key=lambda r: r["ms_per_call"]) if ok else None
    return {"workload": "water-200k-undamped-dipole", "rows": rows, "cheapest_within_1e-6": best}


def target_comparison(dev, steps, plain=((0.3, 8), (0.3, 9), (0.35, 9), (0.35, 10), (0.4, 10)),
                      batched=((0.3, 9), (0.35, 10)), tol=1e-6,
                      sets=32, spots=1024, margins=(0.25, 0.1), safeties=(1.0, 2.0), ncrits=(64,)):
    """Throughput at <= 1e-6 relative RMS field error against the FP64 direct
    sum: the plain FP64 FMM (one multipole set, every pair at expansion_order)
    against the batched (sets on one tree), mixed-precision (FP32 MFMA M2L),
    learned-order FMM.  Workload: the 200,001-site water geometry above,
    undamped (the FMM's far field then has the direct sum's semantics, so the
    error is the method's own), charges plus per-set random dipoles; errors on
    4096 seeded target sites for the first and last set.  The plain side is the
    cheapest of the `plain` (theta, order) settings within tol (the full sweep
    is --plain-sweep); the batched side runs each `batched` (theta, order cap)
    setting at each order margin and order safety factor (the per-pair
    tolerance's factor, FMMH_OPT_ORDER_SAFETY; the spot checks still hold the
    call to tol), and the speedup is the fastest batched run within tol (both
    sets' errors, no fallback) over that plain side.  Both sides run at each
    leaf size of `ncrits` (default: 64, the leaf size of every benchmarked
    configuration).  The selector is trained on the GPU on a 1/20-scale box of
    the same recipe per batched setting and leaf size (the training time is
    reported, not counted in the step)."""
    import torch
    from fmm_amd import synth
    from fmm_amd.octree import Context
    c = synth.config(1)
    x = c["coordinates"]
    n = x.shape[0]
    rng = np.random.default_rng(2024)
    mh = np.zeros((n, 4, sets), order="F")
    for k in range(sets):
        mh[:, 0, k] = c["multipoles"][:, 0]
        mh[:, 1:4, k] = rng.normal(0.0, 0.05, (n, 3))
    X = col(x, dev)
    E = col(c["exclusions"], dev, torch.int32)
    tg = np.sort(np.random.default_rng(4096).choice(n, 4096, replace=False)).astype(np.int32)
    T = torch.as_tensor(tg, device=dev)
    ctx = Context(0)
    ctx.set_timing(True)
    ref = {}
    for k in (0, sets - 1):
        d = col(np.zeros((4096, 4)), dev)
        ctx.field_direct_targets_device(X, col(np.asfortranarray(mh[:, :, k]), dev), E, T, 1, d, "", None)
        ref[k] = d.cpu().numpy()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps

    M1 = col(np.asfortranarray(mh[:, :, 0]), dev)
    o1 = col(np.zeros((n, 4)), dev)
    prow = []
    for nc in ncrits:
        for th_p, p_p in plain:
            def f_plain():
                o1.zero_()
                ctx.field_fmm_device(X, M1, E, th_p, nc, p_p, 1, o1, "", None)
            ms_p = timed(f_plain)
            err_p = rel(o1.cpu().numpy()[tg, 1:4], ref[0][:, 1:4])
            prow.append({"theta": th_p, "expansion_order": p_p, "ncrit": nc, "sets": 1, "ms_per_call": ms_p,
                         "value": n / (ms_p / 1e3), "error": err_p, "within_1e-6": bool(err_p <= tol)})
            print(f"plain theta {th_p} p {p_p} ncrit {nc}: {ms_p:.2f} ms, error {err_p:.3e}", file=sys.stderr,
                  flush=True)
    pok = [r for r in prow if r["within_1e-6"]]
    pbest = min(pok, key=lambda r: r["ms_per_call"]) if pok else None
    MS = col3(mh, dev)
    os_ = col3(np.zeros((n, 4, sets)), dev)
    ct = synth.config(1, scale=0.05)
    mt = np.zeros((ct["coordinates"].shape[0], 4))
    mt[:, 0] = ct["multipoles"][:, 0]
    mt[:, 1:4] = np.random.default_rng(7).normal(0.0, 0.05, (mt.shape[0], 3))
    brow = []
    for nc, (th_b, p_b) in [(nc, b) for nc in ncrits for b in batched]:
        info = ctx.train_selector(ct["coordinates"], np.asfortranarray(mt), th_b, nc, p_b, nsamples=4096, epochs=60,
                                  seed=5)
        ctx.set_precision(True)
        for margin, kap in [(mg, kp) for mg in margins for kp in safeties]:
            ctx.set_order_tolerance(tol, spot_checks=spots, margin=margin, safety=kap)

            def f_batched():
                os_.zero_()
                ctx.field_fmm_sets_device(X, MS, E, th_b, nc, p_b, 1, os_, "", None)
            ms_b = timed(f_batched)
            st = ctx.selection_stats()
            phases = ctx.phase_times()
            ob = os_.cpu().numpy()
            err_b = [rel(ob[tg, 1:4, k], ref[k][:, 1:4]) for k in (0, sets - 1)]
            r = {"theta": th_b, "expansion_order_cap": p_b, "ncrit": nc, "sets": sets, "ms_per_call": ms_b,
                 "value": n * sets / (ms_b / 1e3), "error_set0": err_b[0], "error_last_set": err_b[1],
                 "order_tolerance": tol, "order_margin": margin, "order_safety": kap, "spot_checks": spots,
                 "spot_error": st["spot_error"], "fell_back": st["fallback"], "tolerance_scale": st["tolerance_scale"],
                 "mean_order": st["mean_order"], "m2l_flops_vs_full": st["m2l_flops"] / st["m2l_flops_full"],
                 "phases_ms": {k: round(v, 3) for k, v in phases.items()},
                 "within_1e-6": bool(max(err_b) <= tol and not st["fallback"] and st["tolerance_scale"] == 1.0),
                 "selector_training": info}
            if pbest is not None:
                r["speedup"] = r["value"] / pbest["value"]
            brow.append(r)
            print(f"batched theta {th_b} cap {p_b} ncrit {nc} margin {margin} safety {kap}: {ms_b:.2f} ms, errors "
                  f"{err_b[0]:.3e} {err_b[1]:.3e}, spot {st['spot_error']:.3e}, tolerance scale "
                  f"{st['tolerance_scale']}", file=sys.stderr, flush=True)
        ctx.set_order_tolerance(0.0, safety=1.0)
        ctx.set_precision(False)
    ctx.close()
    bok = [r for r in brow if r["within_1e-6"]]
    bbest = max(bok, key=lambda r: r["value"]) if bok else None
    out = {
        "workload": "water-200k-undamped-dipole (the default water geometry, no damping)",
        "error_metric": "relative RMS of the field block vs FP64 field_direct on 4096 seeded sites",
        "plain_candidates": prow, "plain_fp64": pbest,
        "batched_candidates": brow, "batched_mixed_learned": bbest,
        "speedup": (bbest["value"] / pbest["value"]) if (pbest and bbest) else None,
    }
    return out


def damped_comparison(dev, steps, settings=((0.25, 9), (0.3, 9), (0.3, 10), (0.35, 10)),
                      batched=((0.3, 9), (0.35, 10)), tol=1e-6, spots=1024):
    """The <= 1e-6 comparison on a damped workload: the water geometry with AMOEBA damping and 32 traceless
    quadrupole sets (field order 1).  Errors are against the direct sum with the reference's semantics --
    damping on near-field pairs only (src/fmm.F90:704-736; field_direct damps every pair, :788-815):
    D = D_undamped + (F_damped - F_undamped), F two plain FP64 FMM calls on one tree whose far fields are
    bitwise identical, so their difference is exactly the damping of the near-field pairs.  The plain side is
    the cheapest of `settings` within tol (set 0, 4096 seeded sites); the batched side is the 32 sets in one
    call, mixed precision, learned orders (selector trained per setting on a 1/20-scale box of the same
    recipe), at each of `batched` (theta, order cap); the speedup is the fastest batched setting within tol."""
    import torch
    from fmm_amd import synth
    from fmm_amd.octree import Context
    c = synth.config(2)
    x, df = c["coordinates"], c["damping_factors"]
    n, sets = x.shape[0], c["sets"]
    mh = c["multipoles"]
    X, E = col(x, dev), col(c["exclusions"], dev, torch.int32)
    DF = torch.as_tensor(df, device=dev)
    tg = np.sort(np.random.default_rng(4096).choice(n, 4096, replace=False)).astype(np.int32)
    T = torch.as_tensor(tg, device=dev)
    ctx = Context(0)
    ctx.set_timing(True)
    ref = {}
    for k in (0, sets - 1):
        Mk = col(np.asfortranarray(mh[:, :, k]), dev)
        d = col(np.zeros((4096, 4)), dev)
        ctx.field_direct_targets_device(X, Mk, E, T, 1, d, "", None)
        fd = col(np.zeros((n, 4)), dev)
        ctx.field_fmm_device(X, Mk, E, 0.5, 64, 6, 1, fd, "AMOEBA", DF)
        fu = col(np.zeros((n, 4)), dev)
        ctx.field_fmm_device(X, Mk, E, 0.5, 64, 6, 1, fu, "", None)
        ref[k] = d.cpu().numpy() + (fd.cpu().numpy()[tg] - fu.cpu().numpy()[tg])

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps

    M0 = col(np.asfortranarray(mh[:, :, 0]), dev)
    o1 = col(np.zeros((n, 4)), dev)
    rows = []
    for th, p in settings:
        def f_plain():
            o1.zero_()
            ctx.field_fmm_device(X, M0, E, th, 64, p, 1, o1, "AMOEBA", DF)
        ms = timed(f_plain)
        err = rel(o1.cpu().numpy()[tg, 1:4], ref[0][:, 1:4])
        rows.append({"theta": th, "expansion_order": p, "ms_per_call": ms, "error": err})
        print(f"damped plain theta {th} p {p}: {ms:.2f} ms, error {err:.3e}", file=sys.stderr, flush=True)
    ok = [r for r in rows if r["error"] <= tol]
    best = min(ok, key=lambda r: r["ms_per_call"]) if ok else None
    MS = col3(mh, dev)
    os_ = col3(np.zeros((n, 4, sets)), dev)
    ct = synth.config(2, scale=0.05)
    brows = []
    for th_b, p_b in batched:
        info = ctx.train_selector(ct["coordinates"], np.asfortranarray(ct["multipoles"][:, :, 0]), th_b, 64, p_b,
                                  nsamples=4096, epochs=60, seed=5)
        ctx.set_precision(True)
        ctx.set_order_tolerance(tol, spot_checks=spots)

        def f_batched():
            os_.zero_()
            ctx.field_fmm_sets_device(X, MS, E, th_b, 64, p_b, 1, os_, "AMOEBA", DF)
        ms_b = timed(f_batched)
        st = ctx.selection_stats()
        phases = ctx.phase_times()
        ob = os_.cpu().numpy()
        err_b = [rel(ob[tg, 1:4, k], ref[k][:, 1:4]) for k in (0, sets - 1)]
        ctx.set_order_tolerance(0.0)
        ctx.set_precision(False)
        brows.append({"theta": th_b, "expansion_order_cap": p_b, "sets": sets, "ms_per_call": ms_b,
                      "value": n * sets / (ms_b / 1e3), "error_set0": err_b[0], "error_last_set": err_b[1],
                      "spot_error": st["spot_error"], "fell_back": st["fallback"], "mean_order": st["mean_order"],
                      "within_1e-6": bool(max(err_b) <= tol and not st["fallback"]),
                      "phases_ms": {k: round(v, 3) for k, v in phases.items()}, "selector_training": info})
        print(f"damped batched theta {th_b} cap {p_b}: {ms_b:.2f} ms, errors {err_b[0]:.3e} {err_b[1]:.3e}, "
              f"spot {st['spot_error']:.3e}", file=sys.stderr, flush=True)
    ctx.close()
    bok = [r for r in brows if r["within_1e-6"]]
    bbest = max(bok, key=lambda r: r["value"]) if bok else None
    out = {"workload": "water-200k-amoeba-quadrupole-32sets (field order 1)",
           "error_metric": "relative RMS of the field block vs the direct sum with damping on near-field pairs "
                           "only (the reference's semantics), 4096 seeded sites",
           "plain_sweep": rows, "plain_fp64": best, "batched_sweep": brows, "batched_mixed_learned": bbest}
    if best is not None and bbest is not None:
        best["value"] = n / (best["ms_per_call"] / 1e3)
        out["speedup"] = bbest["value"] / best["value"]
    return out


def bench_uneven(dev, args):
    """Uneven batch: 64 water boxes of 10k-500k sites, 8 multipole sets each, THOLE, field order 1,
    learned M2L orders (cap 10, tolerance 1e-6, 1,024 spot checks per system), mixed precision.  One
    step = one field_fmm_sets call per system, inputs resident in HBM; the selector is trained on a
    1/20-scale box of the default water recipe (time reported, not counted)."""
    import torch
    from fmm_amd import synth
    from fmm_amd.octree import Context
    c = synth.config(4, scale=args.scale)
    th, cap, F, sets = c["theta"], c["expansion_order"], c["field_order"], c["sets"]
    t_start = time.perf_counter()

    def progress(msg):
        print(f"[config 4 {time.perf_counter() - t_start:7.1f} s] {msg}", file=sys.stderr, flush=True)
    progress(f"{len(c['systems'])} systems generated")
    systems = []
    for sy in c["systems"]:
        n = sy["coordinates"].shape[0]
        systems.append((n, col(sy["coordinates"], dev), col3(sy["multipoles"], dev),
                        col(sy["exclusions"], dev, torch.int32), torch.as_tensor(sy["damping_factors"], device=dev),
                        col3(np.zeros((n, synth.ncomp(F), sets)), dev)))
    ctx = Context(0)
    if args.m2l_mfma is not None:
        ctx.set_m2l_mfma(args.m2l_mfma)
    ct = synth.config(1, scale=0.05)
    mt = np.zeros((ct["coordinates"].shape[0], 4))
    mt[:, 0] = ct["multipoles"][:, 0]
    mt[:, 1:4] = np.random.default_rng(7).normal(0.0, 0.05, (mt.shape[0], 3))
    info = ctx.train_selector(ct["coordinates"], np.asfortranarray(mt), th, c["ncrit"], cap, nsamples=4096,
                              epochs=60, seed=5)
    progress("selector trained")
    ctx.set_precision(True)
    ctx.set_order_tolerance(c["order_tolerance"], spot_checks=1024)
    ctx.set_timing(True)
    sel = []

    def step(record=False):
        for k, (n, X, M, E, df, out) in enumerate(systems):
            out.zero_()
            ctx.field_fmm_sets_device(X, M, E, th, c["ncrit"], cap, F, out, c["damp_type"], df)
            if record:
                st = ctx.selection_stats()
                sel.append((n, st, ctx.phase_times()["p2p"], ctx.stats()["p2p_site_pairs"]))
            if k % 16 == 15:
                progress(f"system {k + 1}")

    for w in range(args.warmup):
        step()
        torch.cuda.synchronize()
        progress(f"warm-up step {w + 1} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        torch.cuda.synchronize()
        progress(f"timed step {i + 1} done")
    ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
    step(record=True)
    total = sum(n for n, *_ in systems)
    value = total * sets / (ms_per_step / 1e3)
    fpp = p2p_flops_per_pair(1, F, c["damp_type"], sets)
    pairs = sum(r[3] for r in sel)
    p2p_ms = sum(r[2] for r in sel)
    achieved = pairs * fpp / (p2p_ms / 1e3) / 1e12
    ctx.close()
    return {
        "metric": "site-field evals/s at <=1e-6 rel err vs FP64 direct; % of FP64/FP32 peak",
        "value": value, "unit": "site-field evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": c["name"], "systems": len(systems), "sites": total,
                   "sites_min": min(n for n, *_ in systems), "sites_max": max(n for n, *_ in systems),
                   "sets": sets, "precision": "mixed", "multipole_dim": 4, "field_order": F, "theta": th,
                   "ncrit": c["ncrit"], "expansion_order_cap": cap, "damp_type": c["damp_type"],
                   "order_tolerance": c["order_tolerance"], "spot_checks": 1024},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "kernel": "k_near (P2P, FP64 VALU)",
                     "kernel_ms": p2p_ms, "flop_per_pair": fpp, "site_pairs": pairs,
                     "note": "summed over the 64 calls of one step"},
        "cpu_baseline": None,
        "selection": {"fell_back": sum(int(r[1]["fallback"]) for r in sel),
                      "max_spot_error": max(r[1]["spot_error"] for r in sel),
                      "mean_order": float(np.average([r[1]["mean_order"] for r in sel], weights=[r[0] for r in sel])),
                      "m2l_flops_vs_full": sum(r[1]["m2l_flops"] for r in sel) / sum(r[1]["m2l_flops_full"] for r in sel),
                      "note": "spot error: far field at 1,024 sites per system against the direct far-field sum "
                              "(damping applies to near-field pairs only, as in the reference)"},
        "selector_training": info,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--scale", type=float, default=1.0, help="site-count scale of the uneven batch (config 4)")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default=None,
                    help="default: mixed for config 3 (its workload), fp64 otherwise")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--m2l-mfma", type=int, default=None, choices=[0, 1, 2, 3, 4],
                    help="M2L engine option (FMMH_OPT_M2L_MFMA); default: the library's choice")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (reference for HBM counters) and exit")
    ap.add_argument("--peak-calibration", action="store_true",
                    help="only time large FP32 and FP64 library GEMMs (sustained matrix-core rates) and exit")
    ap.add_argument("--no-target", action="store_true",
                    help="skip the comparison at <= 1e-6 error (plain FP64 vs batched mixed learned-order)")
    ap.add_argument("--target-only", action="store_true", help="only run that comparison and print it")
    ap.add_argument("--target-plain", default="0.3,8;0.3,9;0.35,9;0.35,10;0.4,10",
                    help="theta,p settings of the plain FP64 side (';'-separated; the cheapest within 1e-6 is used)")
    ap.add_argument("--target-batched", default="0.3,9;0.35,10",
                    help="theta,cap settings of the batched side (';'-separated)")
    ap.add_argument("--target-ncrit", default="64",
                    help="leaf sizes of both sides of that comparison (','-separated)")
    ap.add_argument("--target-margin", default="0.25,0.1",
                    help="order margins of the batched side (','-separated)")
    ap.add_argument("--target-safety", default="1,2",
                    help="order safety factors of the batched side (','-separated)")
    ap.add_argument("--damped-target-only", action="store_true",
                    help="only run the <= 1e-6 comparison on the damped quadrupole workload and print it")
    ap.add_argument("--plain-sweep", action="store_true",
                    help="only sweep the plain FP64 side over theta x expansion order (time and error) and exit")
    args = ap.parse_args()
    def settings(arg):
        return tuple((float(a.split(",")[0]), int(a.split(",")[1])) for a in arg.split(";") if a)
    tp, tb_ = settings(args.target_plain), settings(args.target_batched)
    tn = tuple(int(v) for v in args.target_ncrit.split(",") if v)
    tmg = tuple(float(v) for v in args.target_margin.split(",") if v)
    tsf = tuple(float(v) for v in args.target_safety.split(",") if v)
    if args.target_only:
        import torch
        print(json.dumps(target_comparison(torch.device("cuda:0"), args.steps, tp, tb_, ncrits=tn, margins=tmg,
                                           safeties=tsf)))
        return
    if args.damped_target_only:
        import torch
        print(json.dumps(damped_comparison(torch.device("cuda:0"), args.steps)))
        return
    if args.plain_sweep:
        import torch
        print(json.dumps(plain_sweep(torch.device("cuda:0"), args.steps)))
        return
    if args.peak_calibration:
        import torch
        torch.backends.cuda.matmul.allow_tf32 = False
        res = {}
        for name, dt, nsz in (("fp32", torch.float32, 8192), ("fp64", torch.float64, 8192)):
            a = torch.randn(nsz, nsz, dtype=dt, device="cuda:0")
            b = torch.randn(nsz, nsz, dtype=dt, device="cuda:0")
            for _ in range(3):
                c_ = a @ b
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            reps = 10
            for _ in range(reps):
                c_ = a @ b
            torch.cuda.synchronize()
            dt_s = (time.perf_counter() - t0) / reps
            res[name + "_gemm_tflops"] = 2.0 * nsz ** 3 / dt_s / 1e12
        print(json.dumps(res))
        return
    if args.copy_calibration:
        import torch
        src = torch.ones(args.copy_calibration // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel() * 8}))
        return
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)
    if args.config == 4:
        import torch
        print(json.dumps(bench_uneven(torch.device("cuda:0"), args)))
        return

    import torch
    from fmm_amd import synth
    from fmm_amd.octree import Context

    dev = torch.device("cuda:0")
    c = synth.config(args.config)
    n = c["coordinates"].shape[0]
    F = c["field_order"]
    nf = synth.ncomp(F)
    dim = c["multipoles"].shape[1]
    m = {1: 0, 4: 1, 10: 2, 20: 3}[dim]
    X = col(c["coordinates"], dev)
    sets = c["sets"]
    M = col(c["multipoles"], dev) if sets == 1 else col3(c["multipoles"], dev)
    E = col(c["exclusions"], dev, torch.int32)
    df = torch.as_tensor(c["damping_factors"], device=dev) if c["damp_type"] else None
    out = col(np.zeros((n, nf)), dev) if sets == 1 else col3(np.zeros((n, nf, sets)), dev)
    ctx = Context(0)
    ctx.set_timing(True)
    if args.m2l_mfma is not None:
        ctx.set_m2l_mfma(args.m2l_mfma)
    precision = args.precision or ("mixed" if args.config == 3 else "fp64")
    ctx.set_precision(precision == "mixed")

    def step():
        out.zero_()
        torch.cuda.current_stream().synchronize()
        if sets == 1:
            ctx.field_fmm_device(X, M, E, c["theta"], c["ncrit"], c["expansion_order"], F, out, c["damp_type"], df)
        else:
            ctx.field_fmm_sets_device(X, M, E, c["theta"], c["ncrit"], c["expansion_order"], F, out,
                                      c["damp_type"], df)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    phases = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        phases.append(ctx.phase_times())
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    stats = ctx.stats()
    phase_avg = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}

    # Secondary figure: the same call with the tree / list cache on (fixed
    # coordinates, as in induced-dipole iterations).  Not the headline value.
    ctx.set_tree_reuse(True)
    step()
    torch.cuda.synchronize()
    r0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    reuse_ms = (time.perf_counter() - r0) * 1e3 / args.steps
    # and with the device work of those calls replayed as a HIP graph (timing events off: a graph
    # holds no per-phase events); the first call runs eagerly, the second captures
    ctx.set_timing(False)
    ctx.set_graph(True)
    for _ in range(2):
        step()
    torch.cuda.synchronize()
    r0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    graph_ms = (time.perf_counter() - r0) * 1e3 / args.steps
    graph_replayed = ctx.stats()["graph"] == 2
    ctx.set_graph(False)
    ctx.set_timing(True)
    ctx.set_tree_reuse(False)
    value = n * sets / (ms_per_step / 1e3)
    # The drop-in call as a Fortran caller makes it (fmmh_field_fmm: host arrays in, host field out, the
    # exclusion entries checked on the host, pageable copies both ways), on the same inputs
    host_call = None
    if sets == 1:
        hx, hm, he = c["coordinates"], c["multipoles"], c["exclusions"]
        hdf = c["damping_factors"] if c["damp_type"] else ()
        hf = np.zeros((n, nf), order="F")

        def host_step():
            hf[:] = 0.0
            ctx.field_fmm(0, hx, hm, he, c["theta"], c["ncrit"], c["expansion_order"], F, c["damp_type"], hdf,
                          field=hf)
        host_step()
        h0 = time.perf_counter()
        for _ in range(args.steps):
            host_step()
        host_ms = (time.perf_counter() - h0) * 1e3 / args.steps
        hp = ctx.phase_times()
        host_call = {"ms_per_call": host_ms, "value": n / (host_ms / 1e3),
                     "h2d_ms": hp.get("h2d"), "d2h_ms": hp.get("d2h"),
                     "bytes_in": int(hx.nbytes + hm.nbytes + he.nbytes + (hdf.nbytes if c["damp_type"] else 0)
                                     + hf.nbytes), "bytes_out": int(hf.nbytes),
                     "note": "fmmh_field_fmm on host (numpy, column-major) arrays: host check of the exclusion "
                             "entries, pageable copies in (inputs and the field it accumulates into) and out"}

    pairs = stats["p2p_site_pairs"]
    fpp = p2p_flops_per_pair(m, F, c["damp_type"], sets)
    p2p_ms = phase_avg["p2p"]
    achieved = pairs * fpp / (p2p_ms / 1e3) / 1e12
    src = near_src_sha()
    traffic_ok = args.config == 1 and precision == "fp64" and K_NEAR_TRAFFIC_SRC == src
    roofline = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK_TFLOPS,
                "traffic": K_NEAR_TRAFFIC_BYTES if traffic_ok else None,
                "kernel": "k_near (P2P, FP64 VALU)", "kernel_ms": p2p_ms, "flop_per_pair": fpp,
                "site_pairs": pairs, "near_src_sha": src,
                "note": "bound: FP64 compute (the vector-ALU peak equals the FP64 matrix peak on MI355X); "
                        "traffic = HBM bytes per launch recorded by the profile of these near-field sources "
                        "(null when the sources differ from the profiled ones; compulsory ~30 MB)"}

    errors = None
    if not args.no_error and sets > 1:
        # every set against its own FP64 direct sum would cost 32 direct sums: check sets 0 and sets-1
        field = out.cpu().numpy()
        errors = {}
        for k in (0, sets - 1):
            Mk = col(c["multipoles"][:, :, k], dev)
            d = col(np.zeros((n, nf)), dev)
            ctx.field_direct_device(X, Mk, E, F, d, c["damp_type"], df)
            direct = d.cpu().numpy()
            d.zero_()
            torch.cuda.synchronize()
            ctx.field_fmm_device(X, Mk, E, c["theta"], c["ncrit"], c["expansion_order"], F, d, c["damp_type"], df)
            single = d.cpu().numpy()
            errors[f"set{k}_field_block_vs_fp64_direct"] = rel(field[:, 1:4, k], direct[:, 1:4])
            errors[f"set{k}_vs_single_set_fmm"] = rel(field[:, :, k], single)
        errors["note"] = ("damping applies to near-field pairs only, as in the reference, so the damped FMM "
                          "differs from the damped direct sum by the damping of far pairs")
    elif not args.no_error and n > 250_000:
        # full direct sums are O(N^2): state the error on 4096 seeded target sites
        field = out.cpu().numpy()
        tg = np.sort(np.random.default_rng(4096).choice(n, 4096, replace=False)).astype(np.int32)
        T = torch.as_tensor(tg, device=dev)
        d = col(np.zeros((4096, nf)), dev)
        ctx.field_direct_targets_device(X, M, E, T, F, d, c["damp_type"], df)
        direct = d.cpu().numpy()
        d.zero_()
        ctx.field_direct_targets_device(X, M, E, T, F, d, "", None)
        direct_u = d.cpu().numpy()
        u = col(np.zeros((n, nf)), dev)
        ctx.field_fmm_device(X, M, E, c["theta"], c["ncrit"], c["expansion_order"], F, u, "", None)
        fmm_u = u.cpu().numpy()[tg]
        errors = {
            "field_block_vs_fp64_direct": rel(field[tg, 1:4], direct[:, 1:4]),
            "potential_block_vs_fp64_direct": rel(field[tg, :1], direct[:, :1]),
            "field_block_truncation": rel(fmm_u[:, 1:4], direct_u[:, 1:4]),
            "note": "on 4096 seeded target sites; truncation = undamped FMM vs undamped direct",
        }
    elif not args.no_error:
        field = out.cpu().numpy()
        d = col(np.zeros((n, nf)), dev)
        ctx.field_direct_device(X, M, E, F, d, c["damp_type"], df)
        direct = d.cpu().numpy()
        d.zero_()
        torch.cuda.synchronize()
        ctx.field_direct_device(X, M, E, F, d, "", None)
        direct_u = d.cpu().numpy()
        d.zero_()
        torch.cuda.synchronize()
        ctx.field_fmm_device(X, M, E, c["theta"], c["ncrit"], c["expansion_order"], F, d, "", None)
        fmm_u = d.cpu().numpy()
        errors = {
            "field_block_vs_fp64_direct": rel(field[:, 1:4], direct[:, 1:4]),
            "potential_block_vs_fp64_direct": rel(field[:, :1], direct[:, :1]),
            "field_block_truncation": rel(fmm_u[:, 1:4], direct_u[:, 1:4]),
            "note": "truncation = undamped FMM vs undamped direct (damping applies to near-field pairs only, "
                    "as in the reference, so the damped FMM differs from the damped direct sum by the damping "
                    "of far pairs)",
        }
    cpu = None if args.no_cpu else cpu_baseline()
    target = (None if (args.no_target or args.config != 1)
              else target_comparison(dev, min(args.steps, 3), tp, tb_, ncrits=tn, margins=tmg, safeties=tsf))
    damped = None if (args.no_target or args.config != 1) else damped_comparison(dev, min(args.steps, 3))
    line = {
        "metric": "site-field evals/s at <=1e-6 rel err vs FP64 direct; % of FP64/FP32 peak",
        "value": value,
        "value_note": ("site-field evaluations/s of the named workload at its own settings, whose error is "
                       "stated in `errors` (it is not a 1e-6 setting); the throughput at <= 1e-6 is "
                       "`value_at_1e-6`"),
        "value_at_1e-6": (None if target is None else {
            "batched_mixed_learned": target["batched_mixed_learned"] and target["batched_mixed_learned"]["value"],
            "batched_setting": target["batched_mixed_learned"] and {
                k: target["batched_mixed_learned"][k]
                for k in ("theta", "expansion_order_cap", "ncrit", "order_margin", "order_safety")},
            "plain_fp64": target["plain_fp64"] and target["plain_fp64"]["value"],
            "plain_setting": target["plain_fp64"] and {
                k: target["plain_fp64"][k] for k in ("theta", "expansion_order", "ncrit")},
            "leaf_size_note": ("both sides at the leaf size given by --target-ncrit (default 64, that of every "
                               "benchmarked configuration); --target-ncrit 64,80,96 sweeps it on both sides"),
            "speedup": target["speedup"], "workload": target["workload"],
            "damped_speedup": damped and damped.get("speedup"), "damped_workload": damped and damped["workload"]}),
        "unit": "site-field evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if precision == "fp64" else "f64+f32",
        "data": "synthetic",
        "config": {"workload": c["name"], "sites": n, "sets": sets, "precision": precision, "multipole_dim": dim, "field_order": F,
                   "theta": c["theta"], "ncrit": c["ncrit"], "expansion_order": c["expansion_order"],
                   "damp_type": c["damp_type"], "exclusions_per_site": int(c["exclusions"].shape[1])},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "phases_ms": phase_avg,
        "tree_reuse": {"ms_per_step": reuse_ms, "value": n * sets / (reuse_ms / 1e3),
                       "note": "same call with FMMH_OPT_REUSE_TREE on: octree, lists and M2L classes kept from "
                               "the previous call with bitwise-identical coordinates",
                       "graph_ms_per_step": graph_ms, "graph_value": n * sets / (graph_ms / 1e3),
                       "graph_replayed": graph_replayed},
        "drop_in_host_call": host_call,
        "tree": stats,
        "errors": errors,
        "target_1e-6": target,
        "target_1e-6_damped": damped,
    }
    ctx.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
