"""ORACLE -- test infrastructure only.

ctypes front end for
  * liboracle.so: the C++ restatement of peter-reinholdt/fmm's field engine
    (fmm_oracle.cpp), and
  * _ref/libfmmref.so (+ _gnu variant): the reference Fortran sources compiled
    unmodified with the bind(C) wrapper ref_bindc.F90.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package, and only as the checker.  Build with `make -C oracle`.
"""
import ctypes
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
DAMP = {"": 0, "THOLE": 1, "AMOEBA": 2, "ERF": 3}
_libs = {}


def _lib(name):
    if name not in _libs:
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        _libs[name] = ctypes.CDLL(path)
    return _libs[name]


def have_ref(variant=""):
    return os.path.exists(os.path.join(_HERE, "_ref", f"libfmmref{variant}.so"))


_REF_LIBS = {"ref": "_ref/libfmmref.so", "ref_gnu": "_ref/libfmmref_gnu.so", "ref_o3": "_ref/libfmmref_o3.so"}


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def ncomp(k):
    return (k + 1) * (k + 2) * (k + 3) // 6


def _prep(coordinates, multipoles, exclusions, damping_factors):
    X = np.asfortranarray(coordinates, dtype=np.float64)
    M = np.asfortranarray(multipoles, dtype=np.float64)
    E = np.asfortranarray(exclusions, dtype=np.int32)
    if E.ndim == 1:
        E = E.reshape(-1, 1, order="F")
    df = np.ascontiguousarray(damping_factors, dtype=np.float64)
    if df.size == 0:
        df = np.zeros(1)
    return X, M, E, df


def field_direct(coordinates, multipoles, exclusions, field_order, damp_type="", damping_factors=(),
                 impl="oracle", rows=None):
    """impl: 'oracle' (C++ restatement), 'ref' or 'ref_gnu' (compiled reference).
    rows (oracle only): evaluate only these target sites (0-based); returns len(rows) x nf."""
    X, M, E, df = _prep(coordinates, multipoles, exclusions, damping_factors)
    n, dim = M.shape
    d = DAMP[damp_type.strip()]
    if rows is not None:
        assert impl == "oracle", "rows: oracle only"
        R = np.ascontiguousarray(rows, dtype=np.int32)
        F = np.zeros((R.size, ncomp(field_order)), order="F")
        rc = _lib("liboracle.so").orc_field_direct_rows(n, dim, _p(X), _p(M), _p(E), E.shape[1], field_order, d,
                                                        _p(df), _p(R), R.size, _p(F))
        if rc:
            raise ValueError("rows outside 0..n-1")
        return F
    F = np.zeros((n, ncomp(field_order)), order="F")
    if impl == "oracle":
        _lib("liboracle.so").orc_field_direct(n, dim, _p(X), _p(M), _p(E), E.shape[1], field_order, d, _p(df), _p(F))
    else:
        lib = _lib(_REF_LIBS[impl])
        lib.ref_field_direct(n, dim, E.shape[1], _p(X), _p(M), _p(E), field_order, d, _p(df), _p(F))
    return F


def field_fmm(coordinates, multipoles, exclusions, theta, ncrit, expansion_order, field_order, damp_type="",
              damping_factors=(), impl="oracle", box_matrix=None):
    """box_matrix (3 x 3) given: periodic field_fmm, compiled reference only."""
    X, M, E, df = _prep(coordinates, multipoles, exclusions, damping_factors)
    n, dim = M.shape
    F = np.zeros((n, ncomp(field_order)), order="F")
    d = DAMP[damp_type.strip()]
    if box_matrix is not None:
        B = np.asfortranarray(box_matrix, dtype=np.float64)
        if impl == "oracle":
            diag = periodic_diagnostics(X, M, E, theta, ncrit, expansion_order, field_order, damp_type, df, B, F)
            if diag is None:
                raise ValueError("periodic FMM needs n > ncrit and expansion_order >= 2")
            return F
        lib = _lib(_REF_LIBS[impl])
        lib.ref_field_fmm_periodic(n, dim, E.shape[1], _p(X), _p(M), _p(E), ctypes.c_double(theta), ncrit,
                                   expansion_order, field_order, d, _p(df), _p(B), _p(F))
        return F
    if impl == "oracle":
        _lib("liboracle.so").orc_field_fmm(n, dim, _p(X), _p(M), _p(E), E.shape[1], ctypes.c_double(theta), ncrit,
                                           expansion_order, field_order, d, _p(df), _p(F))
    else:
        lib = _lib(_REF_LIBS[impl])
        lib.ref_field_fmm(n, dim, E.shape[1], _p(X), _p(M), _p(E), ctypes.c_double(theta), ncrit, expansion_order,
                          field_order, d, _p(df), _p(F))
    return F


def periodic_diagnostics(X, M, E, theta, ncrit, p, field_order, damp_type, df, box, F, num_nodes=0):
    """Periodic field_fmm of the restatement into F; returns (root local
    expansion after M2L, after the lattice sum, levels summed, the field of
    the near-field pairs alone, all local expansions after L2L when
    num_nodes > 0) or None."""
    n, dim = M.shape
    nm = ncomp(p)
    le1, le2 = np.zeros(nm), np.zeros(nm)
    lv = ctypes.c_int(0)
    fnear = np.zeros_like(F, order="F")
    le_all = np.zeros(max(num_nodes, 1) * nm)
    rc = _lib("liboracle.so").orc_field_fmm_periodic(
        n, dim, _p(X), _p(M), _p(E), E.shape[1], ctypes.c_double(theta), ncrit, p, field_order,
        DAMP[damp_type.strip()], _p(df), _p(box), _p(F), _p(le1), _p(le2), ctypes.byref(lv), _p(fnear),
        _p(le_all) if num_nodes > 0 else None)
    return None if rc else (le1, le2, lv.value, fnear, le_all.reshape(-1, nm) if num_nodes > 0 else None)


def tensor(K, damp_type, x, y, z, a=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    z = np.ascontiguousarray(z, dtype=np.float64)
    n = x.size
    T = np.zeros((n, ncomp(K)), order="F")
    aa = np.ascontiguousarray(a if a is not None else np.zeros(n), dtype=np.float64)
    _lib("liboracle.so").orc_tensor(K, DAMP[damp_type], n, _p(x), _p(y), _p(z), _p(aa), _p(T))
    return T


def norm2(v):
    lib = _lib("liboracle.so")
    lib.orc_norm2.restype = ctypes.c_double
    return lib.orc_norm2(_p(np.ascontiguousarray(v, dtype=np.float64)))


def tree(coordinates, ncrit, theta):
    """Tree and interaction lists of the restatement, pre-order ids, 0-based."""
    lib = _lib("liboracle.so")
    lib.orc_tree_build.restype = ctypes.c_void_p
    X = np.asfortranarray(coordinates, dtype=np.float64)
    n = X.shape[0]
    h = ctypes.c_void_p(lib.orc_tree_build(n, _p(X), ncrit, ctypes.c_double(theta)))
    info = np.zeros(5, dtype=np.int64)
    lib.orc_tree_info(h, _p(info))
    nn, nl, md, nm2l, np2p = (int(v) for v in info)
    out = dict(centers=np.zeros((nn, 3), order="F"), r=np.zeros(nn), rmax=np.zeros(nn),
               depth=np.zeros(nn, np.int32), parent=np.zeros(nn, np.int32), is_leaf=np.zeros(nn, np.int32),
               nleaf=np.zeros(nn, np.int32), mem_off=np.zeros(nn + 1, np.int32), members=np.zeros(n, np.int32),
               m2l_off=np.zeros(nn + 1, np.int32), m2l_src=np.zeros(max(nm2l, 1), np.int32),
               p2p_off=np.zeros(nn + 1, np.int32), p2p_src=np.zeros(max(np2p, 1), np.int32))
    lib.orc_tree_export(h, *[_p(out[k]) for k in ("centers", "r", "rmax", "depth", "parent", "is_leaf", "nleaf",
                                                   "mem_off", "members", "m2l_off", "m2l_src", "p2p_off",
                                                   "p2p_src")])
    lib.orc_tree_free(h)
    out["m2l_src"] = out["m2l_src"][:nm2l]
    out["p2p_src"] = out["p2p_src"][:np2p]
    out["max_depth"] = md
    out["num_leaves"] = nl
    return out


# ---------------------------------------------------------------- QM/MM embedding
def get_boxes(coordinates, target_coordinates, theta, ncrit, expansion_order, impl="ref"):
    """(idx_near_field, box_coordinates, box_indices), 1-based as the reference returns them."""
    X = np.asfortranarray(coordinates, dtype=np.float64)
    T = np.asfortranarray(target_coordinates, dtype=np.float64)
    n, nt = X.shape[0], T.shape[0]
    sizes = np.zeros(2, np.int32)
    near = np.zeros(n, np.int32)
    nodes = 8 * n + 8
    bc = np.zeros((nodes, 3), order="F")
    bi = np.zeros(nodes, np.int32)
    if impl == "oracle":
        _lib("liboracle.so").orc_get_boxes(n, _p(X), nt, _p(T), ctypes.c_double(theta), ncrit, nodes, _p(sizes),
                                           _p(near), _p(bc), _p(bi))
    else:
        _lib("_ref/libfmmref.so").ref_get_boxes(n, _p(X), nt, _p(T), ctypes.c_double(theta), ncrit, expansion_order,
                                                n, nodes, _p(sizes), _p(near), _p(bc), _p(bi))
    return near[: sizes[0]].copy(), np.asfortranarray(bc[: sizes[1]]), bi[: sizes[1]].copy()


def get_multipoles(coordinates, multipoles, target_coordinates, theta, ncrit, expansion_order, impl="ref"):
    """(idx_near_field, box_coordinates, box_multipoles); multipoles N x dim x ndens."""
    X = np.asfortranarray(coordinates, dtype=np.float64)
    M = np.asfortranarray(multipoles, dtype=np.float64)
    T = np.asfortranarray(target_coordinates, dtype=np.float64)
    n, dim, nd = M.shape
    nt = T.shape[0]
    nm = ncomp(expansion_order)
    sizes = np.zeros(2, np.int32)
    near = np.zeros(n, np.int32)
    nodes = 8 * n + 8
    bc = np.zeros((nodes, 3), order="F")
    bm = np.zeros((nodes, nm, nd), order="F")
    if impl == "oracle":
        _lib("liboracle.so").orc_get_multipoles(n, dim, nd, _p(X), _p(M), nt, _p(T), ctypes.c_double(theta), ncrit,
                                                expansion_order, nodes, _p(sizes), _p(near), _p(bc), _p(bm))
    else:
        _lib("_ref/libfmmref.so").ref_get_multipoles(n, dim, nd, _p(X), _p(M), nt, _p(T), ctypes.c_double(theta),
                                                     ncrit, expansion_order, n, nodes, _p(sizes), _p(near), _p(bc),
                                                     _p(bm))
    return near[: sizes[0]].copy(), np.asfortranarray(bc[: sizes[1]]), np.asfortranarray(bm[: sizes[1]])


def interpolate_electron_fields(coordinates, local_expansion, box_indices, theta, ncrit, expansion_order,
                                field_order, impl="ref", out=None):
    """impl='ref' aborts in the reference's finalize after the field is
    written (see tests/golden/make_golden.py); pass `out` in shared memory."""
    X = np.asfortranarray(coordinates, dtype=np.float64)
    L = np.asfortranarray(local_expansion, dtype=np.float64)
    B = np.ascontiguousarray(box_indices, dtype=np.int32)
    n = X.shape[0]
    nb, nm, nd = L.shape
    F = np.zeros((n, ncomp(field_order), nd), order="F") if out is None else out
    if impl == "oracle":
        _lib("liboracle.so").orc_interpolate_electron_fields(n, _p(X), nb, nm, nd, _p(L), _p(B),
                                                             ctypes.c_double(theta), ncrit, expansion_order,
                                                             field_order, _p(F))
    else:
        _lib("_ref/libfmmref.so").ref_interpolate_electron_fields(n, _p(X), nb, nm, nd, _p(L), _p(B),
                                                                  ctypes.c_double(theta), ncrit, expansion_order,
                                                                  field_order, _p(F))
    return F
