"""Write the golden fixtures of tests/golden/ from the reference itself.

The reference Fortran (peter-reinholdt/fmm src/*.F90) is compiled unmodified
by oracle/Makefile into oracle/_ref/ (flang, the shipped -O1 -g flags); this
script calls its field_direct / field_fmm through the bind(C) wrapper
oracle/ref_bindc.F90 and stores inputs and outputs only.

Run in the build container (it needs the reference sources):
    make -C oracle ref && python tests/golden/make_golden.py [--periodic | --qmmm]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from fmm_amd import synth  # noqa: E402


def water_inputs(n_mol, dim, seed):
    x, q, excl, df, rng = synth.water_box(n_mol, seed)
    n = x.shape[0]
    m = np.zeros((n, dim))
    m[:, 0] = q
    if dim > 1:
        m[:, 1:4] = rng.normal(0.0, 0.05, (n, 3))
    if dim > 4:
        m[:, 4:] = rng.normal(0.0, 0.05, (n, dim - 4))
    synth.traceless(m)
    erf_df = rng.uniform(0.8, 1.2, n)
    return x, np.asfortranarray(m), excl, df, erf_df


def main():
    assert oracle.have_ref() and oracle.have_ref("_gnu"), "build oracle/_ref first (make -C oracle ref)"
    # 1. the CPU-runnable benchmark workload: 4096 uniform sites, quadrupoles
    c = synth.config(0)
    X, M, E = c["coordinates"], c["multipoles"], c["exclusions"]
    fd = oracle.field_direct(X, M, E, 1, impl="ref")
    ff = oracle.field_fmm(X, M, E, 0.5, 64, 5, 1, impl="ref")
    np.savez_compressed(os.path.join(HERE, "uniform4096_p5.npz"), coordinates=X, multipoles=M, exclusions=E,
                        theta=0.5, ncrit=64, expansion_order=5, field_order=1,
                        field_direct=fd, field_fmm=ff)
    # 2. 999-site water box, quadrupoles, every damping kind, field order 2
    x, m, excl, df, erf_df = water_inputs(333, 10, 7)
    out = dict(coordinates=x, multipoles=m, exclusions=excl, thole_df=df, erf_df=erf_df, theta=0.5, ncrit=64,
               expansion_order=6, field_order=2)
    for damp in ("", "THOLE", "AMOEBA", "ERF"):
        d = erf_df if damp == "ERF" else df
        key = damp.lower() or "none"
        out[f"direct_{key}"] = oracle.field_direct(x, m, excl, 2, damp, d, impl="ref")
        out[f"fmm_{key}"] = oracle.field_fmm(x, m, excl, 0.5, 64, 6, 2, damp, d, impl="ref")
    np.savez_compressed(os.path.join(HERE, "water999_q.npz"), **out)
    # 3. expansion-order sweep (undamped dipoles) incl. the orders where the
    #    reference's int32 T-tensor coefficients overflow (p >= 7 with flang,
    #    p >= 6 on the gfortran branch), and theta 0.6 / ncrit 16 variants
    x, m, excl, df, _ = water_inputs(333, 4, 11)
    out = dict(coordinates=x, multipoles=m, exclusions=excl)
    out["direct"] = oracle.field_direct(x, m, excl, 1, impl="ref")
    for p in (2, 3, 4, 5, 6, 7, 8):
        out[f"fmm_p{p}"] = oracle.field_fmm(x, m, excl, 0.5, 64, p, 1, impl="ref")
        out[f"fmm_gnu_p{p}"] = oracle.field_fmm(x, m, excl, 0.5, 64, p, 1, impl="ref_gnu")
    out["fmm_theta06_ncrit16_p5"] = oracle.field_fmm(x, m, excl, 0.6, 16, 5, 1, impl="ref")
    np.savez_compressed(os.path.join(HERE, "water999_dipole_orders.npz"), **out)
    # 4. charges only, N <= ncrit fallback and a tiny tree
    rng = np.random.default_rng(5)
    for n, name in ((40, "tiny40"), (300, "small300")):
        x = np.asfortranarray(rng.uniform(0, 10, (n, 3)))
        q = rng.normal(0, 1, (n, 1))
        q -= q.mean()
        q = np.asfortranarray(q)
        e = synth.self_exclusions(n)
        np.savez_compressed(os.path.join(HERE, f"charges_{name}.npz"), coordinates=x, multipoles=q, exclusions=e,
                            field_direct=oracle.field_direct(x, q, e, 2, impl="ref"),
                            field_fmm=oracle.field_fmm(x, q, e, 0.5, 64, 4, 2, impl="ref"))
    periodic()
    qmmm()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


def periodic():
    """5. periodic field_fmm (src/fmm.F90:1392-1395) where the reference's
    lattice sum converges: the 999-site water box (a polar cell) in an exactly
    cubic box, whose cubic shells cancel the dipole term, and the same box
    plus its inversion image (no net dipole) in a triclinic box.  In other
    boxes the reference's periodic_sum runs all 50 levels without converging
    and its output is rounding noise (the potential entirely, the field at
    p >= 6), so no implementation can be held to it there; every case written
    here is checked to agree between the two reference builds."""
    x, m, excl, df, erf_df = water_inputs(333, 4, 13)
    L = float((x.max(axis=0) - x.min(axis=0)).max()) + 2.0
    cube = np.asfortranarray(np.diag([L, L, L]))
    c = x.mean(axis=0)
    n = x.shape[0]
    xs = np.asfortranarray(np.vstack([x, 2.0 * c - x]))
    ms = np.asfortranarray(np.vstack([m, m * np.array([1.0, -1.0, -1.0, -1.0])]))
    es = np.asfortranarray(np.vstack([excl, np.where(excl > 0, excl + n, 0)]).astype(np.int32))
    dfs = np.concatenate([df, df])
    erf_dfs = np.concatenate([erf_df, erf_df])
    Ls = float((xs.max(axis=0) - xs.min(axis=0)).max()) + 2.0
    tric = np.asfortranarray(np.diag([Ls, Ls, Ls]))
    tric[0, 1] = 3.0          # second box vector tilted along x
    tric[0, 2] = -2.0         # third box vector tilted along x and y
    tric[1, 2] = 1.5
    out = dict(coordinates=x, multipoles=m, exclusions=excl, thole_df=df, erf_df=erf_df, box_cube=cube,
               sym_coordinates=xs, sym_multipoles=ms, sym_exclusions=es, sym_thole_df=dfs, sym_erf_df=erf_dfs,
               box_triclinic=tric)
    cases = {"cube_thole_p6_f2": ("", cube, "THOLE", 6, 2), "cube_none_p4_f1": ("", cube, "", 4, 1),
             "cube_amoeba_p5_f1": ("", cube, "AMOEBA", 5, 1), "cube_erf_p5_f2": ("", cube, "ERF", 5, 2),
             "sym_triclinic_none_p5_f1": ("sym_", tric, "", 5, 1),
             "sym_triclinic_thole_p6_f2": ("sym_", tric, "THOLE", 6, 2)}
    for key, (pre, box, damp, p, F) in cases.items():
        d = out[pre + ("erf_df" if damp == "ERF" else "thole_df")]
        args = (out[pre + "coordinates"], out[pre + "multipoles"], out[pre + "exclusions"], 0.5, 64, p, F, damp, d)
        ref = oracle.field_fmm(*args, impl="ref", box_matrix=box)
        # the lattice sum must converge before its 50-level limit ...
        X, M, E = (np.asfortranarray(a) for a in args[:3])
        levels = oracle.periodic_diagnostics(X, M, E.astype(np.int32), 0.5, 64, p, F, damp,
                                             np.ascontiguousarray(d, dtype=np.float64), box,
                                             np.zeros_like(ref, order="F"))[2]
        assert levels < 50, (key, levels)
        # ... and, where the gfortran branch is valid (its recursive tensor
        # coefficients overflow int32 from order 12, i.e. p >= 6), both builds agree
        if p <= 5:
            gnu = oracle.field_fmm(*args, impl="ref_gnu", box_matrix=box)
            spread = max(np.linalg.norm(ref[:, lo:hi] - gnu[:, lo:hi]) / np.linalg.norm(ref[:, lo:hi])
                         for lo, hi in ((0, 1), (1, 4), (4, 10))[: F + 1])
            assert spread < 1e-13, (key, spread)
        out[key] = ref
        out[key + "_levels"] = levels
    np.savez_compressed(os.path.join(HERE, "water999_periodic.npz"), **out)


def _ref_interpolate(x, le, bi, theta, ncrit, p, F):
    """The reference's interpolate_electron_fields writes the field and then
    aborts in finalize: clean_node deallocates interaction lists that the
    QM/MM routines never allocate (src/fmm.F90:318-321, 1335).  Run it in a
    forked child that writes into shared memory and read the field back."""
    from multiprocessing import shared_memory
    shape = (x.shape[0], (F + 1) * (F + 2) * (F + 3) // 6, le.shape[2])
    shm = shared_memory.SharedMemory(create=True, size=int(np.prod(shape)) * 8)
    try:
        out = np.ndarray(shape, dtype=np.float64, buffer=shm.buf, order="F")
        out[:] = 0.0
        pid = os.fork()
        if pid == 0:
            os.close(2)                     # the runtime's abort message
            oracle.interpolate_electron_fields(x, le, bi, theta, ncrit, p, F, impl="ref", out=out)
            os._exit(0)
        os.waitpid(pid, 0)
        return np.array(out, order="F")
    finally:
        shm.close()
        shm.unlink()


def qmmm():
    """6. QM/MM embedding (src/fmm.F90:1134-1336): the 999-site water box as
    the MM region and a few QM target points around its middle.  The
    reference's get_boxes / get_multipoles abort before returning (see
    _ref_interpolate), so their outputs here come from the restatement and
    serve as inputs; the pinned output is the reference's own
    interpolate_electron_fields on those boxes."""
    x, m, excl, df, _ = water_inputs(333, 4, 17)
    rng = np.random.default_rng(17)
    c = x.mean(axis=0)
    targets = np.asfortranarray(c + rng.normal(0.0, 2.0, (6, 3)))
    m2 = np.asfortranarray(np.stack([m, rng.normal(0.0, 0.1, m.shape)], axis=2))
    out = dict(coordinates=x, multipoles=m2, targets=targets)
    for theta, ncrit, p, F in ((0.5, 64, 4, 1), (0.3, 16, 6, 2)):
        key = f"t{int(theta * 10)}_n{ncrit}_p{p}"
        near, bc, bi = oracle.get_boxes(x, targets, theta, ncrit, p, impl="oracle")
        le = np.asfortranarray(rng.normal(0.0, 1e-3, (len(bi), (p + 1) * (p + 2) * (p + 3) // 6, 2)))
        ref = _ref_interpolate(x, le, bi, theta, ncrit, p, F)
        assert np.abs(ref).max() > 0 and np.isfinite(ref).all()
        out.update({f"{key}_box_indices": bi, f"{key}_local_expansion": le, f"{key}_field": ref,
                    f"{key}_field_order": F})
    np.savez_compressed(os.path.join(HERE, "water999_qmmm.npz"), **out)


if __name__ == "__main__":
    assert oracle.have_ref(), "build oracle/_ref first (make -C oracle ref)"
    if sys.argv[1:] == ["--periodic"]:
        periodic()
    elif sys.argv[1:] == ["--qmmm"]:
        qmmm()
    else:
        main()
