"""Python face of the field engine, mirroring the f2py module `octree` that
peter-reinholdt/fmm builds (CMakeLists.txt:21-39, `only: field_fmm
field_direct`).  As with f2py, the intent(out) `field` is returned rather
than passed, and arrays may be any numpy layout (they are copied to
Fortran order).

    field = field_direct(comm, coordinates, multipoles, exclusions, field_order,
                         damp_type, damping_factors)
    field = field_fmm(comm, coordinates, multipoles, exclusions, theta, ncrit,
                      expansion_order, field_order, damp_type, damping_factors,
                      is_periodic, box_matrix)

Errors raise FmmError carrying the reference's messages ("Unknown damp
type: ...", src/fmm.F90:594, 1376).  All compute runs in libfmm_hip.so on the
GPU; there is no CPU path.

`Context` adds what f2py cannot express: a dedicated stream with pooled
device memory, device-resident calls on torch tensors (`*_device`), phase
timers and tree export.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import FmmError, check

__all__ = ["field_direct", "field_fmm", "get_multipoles", "get_boxes", "interpolate_electron_fields", "Context",
           "FmmError", "ncomp"]

DAMP = {"": 0, "THOLE": 1, "AMOEBA": 2, "ERF": 3}
PHASE_CAP = 64


def ncomp(k):
    return (k + 1) * (k + 2) * (k + 3) // 6


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _damp_code(damp_type):
    lib = _lib.load()
    raw = damp_type.encode() if isinstance(damp_type, str) else bytes(damp_type)
    code = ctypes.c_int32(0)
    check(lib.fmmh_parse_damp_type(raw, len(raw), ctypes.byref(code)))
    return code.value


_BOXES = []


def _periodic(is_periodic, box_matrix):
    """(flag, host pointer) for the C ABI; the box stays alive for the call."""
    if not is_periodic:
        return 0, None
    box = np.asfortranarray(box_matrix, dtype=np.float64)
    if box.shape != (3, 3):
        raise ValueError("box_matrix must be 3 x 3")
    _BOXES[:] = [box]
    return 1, _p(box)


def _prep(coordinates, multipoles, exclusions, damping_factors):
    X = np.asfortranarray(coordinates, dtype=np.float64)
    M = np.asfortranarray(multipoles, dtype=np.float64)
    E = np.asfortranarray(exclusions, dtype=np.int32)
    if E.ndim == 1:
        E = E.reshape(-1, 1, order="F")
    if X.ndim != 2 or X.shape[1] != 3:
        raise ValueError("coordinates must be N x 3")
    if M.ndim != 2 or M.shape[0] != X.shape[0]:
        raise ValueError("multipoles must be N x dim")
    if E.shape[0] != X.shape[0]:
        raise ValueError("exclusions must be N x nexcl")
    df = np.ascontiguousarray(damping_factors, dtype=np.float64).reshape(-1)
    return X, M, E, df


class Context:
    """One HIP stream plus pooled device buffers on one device."""

    def __init__(self, device=-1):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        check(self._lib.fmmh_context_create(ctypes.byref(h), device))
        self._h = h

    def close(self):
        if self._h:
            check(self._lib.fmmh_context_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # -- host arrays ---------------------------------------------------------
    def field_direct(self, comm, coordinates, multipoles, exclusions, field_order, damp_type="",
                     damping_factors=(), field=None):
        X, M, E, df = _prep(coordinates, multipoles, exclusions, damping_factors)
        n, dim = M.shape
        code = _damp_code(damp_type)
        F = np.zeros((n, ncomp(field_order)), order="F") if field is None else field
        check(self._lib.fmmh_field_direct(self._h, int(comm), _p(X), n, _p(M), n, n, dim, _p(E), n, E.shape[1],
                                          int(field_order), _p(F), F.shape[0], code,
                                          _p(df) if df.size else None, df.size))
        return F

    def field_fmm(self, comm, coordinates, multipoles, exclusions, theta, ncrit, expansion_order, field_order,
                  damp_type="", damping_factors=(), is_periodic=False, box_matrix=None, field=None):
        X, M, E, df = _prep(coordinates, multipoles, exclusions, damping_factors)
        n, dim = M.shape
        code = _damp_code(damp_type)
        F = np.zeros((n, ncomp(field_order)), order="F") if field is None else field
        box = np.zeros((3, 3), order="F") if box_matrix is None else np.asfortranarray(box_matrix, np.float64)
        check(self._lib.fmmh_field_fmm(self._h, int(comm), _p(X), n, _p(M), n, n, dim, _p(E), n, E.shape[1],
                                       float(theta), int(ncrit), int(expansion_order), int(field_order), _p(F),
                                       F.shape[0], code, _p(df) if df.size else None, df.size,
                                       1 if is_periodic else 0, _p(box)))
        return F

    # -- many multipole sets on one set of sites --------------------------------
    @staticmethod
    def _prep_sets(coordinates, multipoles, exclusions, damping_factors):
        M = np.asfortranarray(multipoles, dtype=np.float64)
        if M.ndim == 2:
            M = M.reshape(M.shape[0], M.shape[1], 1, order="F")
        if M.ndim != 3:
            raise ValueError("multipoles must be N x dim x nsets")
        X, _, E, df = _prep(coordinates, M.reshape(M.shape[0], -1, order="F"), exclusions, damping_factors)
        return X, M, E, df

    def field_direct_sets(self, comm, coordinates, multipoles, exclusions, field_order, damp_type="",
                          damping_factors=(), field=None):
        """field_direct for nsets multipole sets (multipoles N x dim x nsets); returns N x nf x nsets."""
        X, M, E, df = self._prep_sets(coordinates, multipoles, exclusions, damping_factors)
        n, dim, ns = M.shape
        F = np.zeros((n, ncomp(field_order), ns), order="F") if field is None else field
        check(self._lib.fmmh_field_direct_sets(self._h, int(comm), _p(X), n, _p(M), n, n, dim, ns, _p(E), n,
                                               E.shape[1], int(field_order), _p(F), F.shape[0],
                                               _damp_code(damp_type), _p(df) if df.size else None, df.size))
        return F

    def field_fmm_sets(self, comm, coordinates, multipoles, exclusions, theta, ncrit, expansion_order, field_order,
                       damp_type="", damping_factors=(), is_periodic=False, box_matrix=None, field=None):
        """field_fmm for nsets multipole sets on one tree (multipoles N x dim x nsets); returns N x nf x nsets."""
        X, M, E, df = self._prep_sets(coordinates, multipoles, exclusions, damping_factors)
        n, dim, ns = M.shape
        F = np.zeros((n, ncomp(field_order), ns), order="F") if field is None else field
        box = np.zeros((3, 3), order="F") if box_matrix is None else np.asfortranarray(box_matrix, np.float64)
        check(self._lib.fmmh_field_fmm_sets(self._h, int(comm), _p(X), n, _p(M), n, n, dim, ns, _p(E), n,
                                            E.shape[1], float(theta), int(ncrit), int(expansion_order),
                                            int(field_order), _p(F), F.shape[0], _damp_code(damp_type),
                                            _p(df) if df.size else None, df.size, 1 if is_periodic else 0, _p(box)))
        return F

    # -- QM/MM embedding (src/fmm.F90:1134-1336) --------------------------------
    def get_boxes(self, coordinates, target_coordinates, theta, ncrit, expansion_order):
        """(idx_near_field, box_coordinates, box_indices), indices 1-based as in the reference."""
        X = np.asfortranarray(coordinates, dtype=np.float64)
        T = np.asfortranarray(target_coordinates, dtype=np.float64).reshape(-1, 3, order="F")
        sizes = np.zeros(2, np.int64)
        check(self._lib.fmmh_get_boxes(self._h, _p(X), X.shape[0], X.shape[0], _p(T), max(T.shape[0], 1),
                                       T.shape[0], float(theta), int(ncrit), int(expansion_order), _p(sizes)))
        near = np.zeros(max(sizes[0], 1), np.int32)
        boxes = np.zeros((max(sizes[1], 1), 3), order="F")
        idx = np.zeros(max(sizes[1], 1), np.int32)
        check(self._lib.fmmh_get_boxes_export(self._h, _p(near), _p(boxes), _p(idx)))
        return near[: sizes[0]], np.asfortranarray(boxes[: sizes[1]]), idx[: sizes[1]]

    def get_multipoles(self, comm, coordinates, multipoles, target_coordinates, theta, ncrit, expansion_order):
        """(idx_near_field, box_coordinates, box_multipoles); multipoles N x dim x ndens."""
        X = np.asfortranarray(coordinates, dtype=np.float64)
        M = np.asfortranarray(multipoles, dtype=np.float64)
        if M.ndim == 2:
            M = M.reshape(M.shape[0], M.shape[1], 1, order="F")
        T = np.asfortranarray(target_coordinates, dtype=np.float64).reshape(-1, 3, order="F")
        n, dim, nd = M.shape
        sizes = np.zeros(2, np.int64)
        check(self._lib.fmmh_get_multipoles(self._h, int(comm), _p(X), n, n, _p(M), n, dim, nd, _p(T),
                                            max(T.shape[0], 1), T.shape[0], float(theta), int(ncrit),
                                            int(expansion_order), _p(sizes)))
        nf_, nb = int(sizes[0]), int(sizes[1])
        near = np.zeros(max(nf_, 1), np.int32)
        boxes = np.zeros((max(nb, 1), 3), order="F")
        bm = np.zeros((max(nb, 1), ncomp(expansion_order), nd), order="F")
        check(self._lib.fmmh_get_multipoles_export(self._h, _p(near), _p(boxes), _p(bm)))
        return near[:nf_], np.asfortranarray(boxes[:nb]), np.asfortranarray(bm[:nb])

    def interpolate_electron_fields(self, comm, coordinates, local_expansion, box_indices, theta, ncrit,
                                    expansion_order, field_order, field=None):
        """Field (N x nf x ndens) of local expansions placed on the given boxes (1-based indices)."""
        X = np.asfortranarray(coordinates, dtype=np.float64)
        L = np.asfortranarray(local_expansion, dtype=np.float64)
        if L.ndim == 2:
            L = L.reshape(L.shape[0], L.shape[1], 1, order="F")
        B = np.ascontiguousarray(box_indices, dtype=np.int32)
        n = X.shape[0]
        nb, nm, nd = L.shape
        if nm != ncomp(expansion_order) or B.size != nb:
            raise ValueError("local_expansion must be nbox x ncomp(expansion_order) x ndens, one index per box")
        F = np.zeros((n, ncomp(field_order), nd), order="F") if field is None else field
        check(self._lib.fmmh_interpolate_electron_fields(self._h, int(comm), _p(X), n, n, _p(L), max(nb, 1), nb, nd,
                                                         _p(B) if nb else None, float(theta), int(ncrit),
                                                         int(expansion_order), int(field_order), _p(F), F.shape[0]))
        return F

    # -- device arrays (torch tensors already in HBM) --------------------------
    @staticmethod
    def _dptr(t):
        return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() else None

    @staticmethod
    def _ld(t):
        """Leading dimension of a column-major (Fortran-order) 2-D tensor."""
        if t.dim() == 1:
            return t.shape[0]
        if t.stride(0) != 1:
            raise ValueError("device arrays must be column-major: use t.T.contiguous().T")
        return t.stride(1) if t.shape[1] > 1 else t.shape[0]

    def field_direct_device(self, coordinates, multipoles, exclusions, field_order, field, damp_type="",
                            damping_factors=None):
        n, dim = multipoles.shape
        nexcl = exclusions.shape[1] if exclusions.dim() == 2 else 1
        ndf = damping_factors.numel() if damping_factors is not None else 0
        check(self._lib.fmmh_field_direct_dev(
            self._h, 0, self._dptr(coordinates), self._ld(coordinates), self._dptr(multipoles),
            self._ld(multipoles), n, dim, self._dptr(exclusions), self._ld(exclusions), nexcl, int(field_order),
            self._dptr(field), self._ld(field), _damp_code(damp_type), self._dptr(damping_factors), ndf))
        return field

    def field_direct_device_noalloc(self, coordinates, multipoles, exclusions, field_order, field, damp_type="",
                                    damping_factors=None):
        """field_direct_device with the working arrays in the library's fixed per-device memory: no
        allocation (fmmh_field_direct_dev_noalloc)."""
        n, dim = multipoles.shape
        nexcl = exclusions.shape[1] if exclusions.dim() == 2 else 1
        ndf = damping_factors.numel() if damping_factors is not None else 0
        check(self._lib.fmmh_field_direct_dev_noalloc(
            self._h, 0, self._dptr(coordinates), self._ld(coordinates), self._dptr(multipoles),
            self._ld(multipoles), n, dim, self._dptr(exclusions), self._ld(exclusions), nexcl, int(field_order),
            self._dptr(field), self._ld(field), _damp_code(damp_type), self._dptr(damping_factors), ndf))
        return field

    def field_fmm_device(self, coordinates, multipoles, exclusions, theta, ncrit, expansion_order, field_order,
                         field, damp_type="", damping_factors=None, is_periodic=False, box_matrix=None):
        n, dim = multipoles.shape
        nexcl = exclusions.shape[1] if exclusions.dim() == 2 else 1
        ndf = damping_factors.numel() if damping_factors is not None else 0
        check(self._lib.fmmh_field_fmm_dev(
            self._h, 0, self._dptr(coordinates), self._ld(coordinates), self._dptr(multipoles),
            self._ld(multipoles), n, dim, self._dptr(exclusions), self._ld(exclusions), nexcl, float(theta),
            int(ncrit), int(expansion_order), int(field_order), self._dptr(field), self._ld(field),
            _damp_code(damp_type), self._dptr(damping_factors), ndf, *_periodic(is_periodic, box_matrix)))
        return field

    @staticmethod
    def _ld3(t):
        """Leading dimension of a column-major N x c x nsets tensor."""
        if t.stride(0) != 1 or t.stride(2) != t.stride(1) * t.shape[1]:
            raise ValueError("device set arrays must be column-major N x c x nsets: use t.permute(2, 1, 0)"
                             ".contiguous().permute(2, 1, 0)")
        return t.stride(1)

    def field_direct_sets_device(self, coordinates, multipoles, exclusions, field_order, field, damp_type="",
                                 damping_factors=None):
        n, dim, ns = multipoles.shape
        nexcl = exclusions.shape[1] if exclusions.dim() == 2 else 1
        ndf = damping_factors.numel() if damping_factors is not None else 0
        check(self._lib.fmmh_field_direct_sets_dev(
            self._h, 0, self._dptr(coordinates), self._ld(coordinates), self._dptr(multipoles),
            self._ld3(multipoles), n, dim, ns, self._dptr(exclusions), self._ld(exclusions), nexcl,
            int(field_order), self._dptr(field), self._ld3(field), _damp_code(damp_type),
            self._dptr(damping_factors), ndf))
        return field

    def field_fmm_sets_device(self, coordinates, multipoles, exclusions, theta, ncrit, expansion_order,
                              field_order, field, damp_type="", damping_factors=None, is_periodic=False,
                              box_matrix=None):
        n, dim, ns = multipoles.shape
        nexcl = exclusions.shape[1] if exclusions.dim() == 2 else 1
        ndf = damping_factors.numel() if damping_factors is not None else 0
        check(self._lib.fmmh_field_fmm_sets_dev(
            self._h, 0, self._dptr(coordinates), self._ld(coordinates), self._dptr(multipoles),
            self._ld3(multipoles), n, dim, ns, self._dptr(exclusions), self._ld(exclusions), nexcl, float(theta),
            int(ncrit), int(expansion_order), int(field_order), self._dptr(field), self._ld3(field),
            _damp_code(damp_type), self._dptr(damping_factors), ndf, *_periodic(is_periodic, box_matrix)))
        return field

    def field_direct_targets_device(self, coordinates, multipoles, exclusions, targets, field_order, field,
                                    damp_type="", damping_factors=None):
        """Direct sum at the sites `targets` (0-based int32 device tensor); field is ntargets x nf."""
        n, dim = multipoles.shape
        nexcl = exclusions.shape[1] if exclusions.dim() == 2 else 1
        ndf = damping_factors.numel() if damping_factors is not None else 0
        check(self._lib.fmmh_field_direct_targets_dev(
            self._h, self._dptr(coordinates), self._ld(coordinates), self._dptr(multipoles), self._ld(multipoles),
            n, dim, self._dptr(exclusions), self._ld(exclusions), nexcl, self._dptr(targets), targets.numel(),
            int(field_order), self._dptr(field), self._ld(field), _damp_code(damp_type),
            self._dptr(damping_factors), ndf))
        return field

    # -- diagnostics -----------------------------------------------------------
    def set_timing(self, on=True):
        check(self._lib.fmmh_set_option(self._h, 1, 1.0 if on else 0.0))

    def set_list_capacity(self, scale=1.0):
        """Scale of the initial interaction-list buffers (they grow on overflow)."""
        check(self._lib.fmmh_set_option(self._h, 2, float(scale)))

    def set_precision(self, mixed=False):
        """False: FP64 throughout (default).  True: M2L translations on FP32 MFMA (FP64 near field, P2M,
        L2P and accumulation)."""
        check(self._lib.fmmh_set_option(self._h, 5, 1.0 if mixed else 0.0))

    def set_tree_reuse(self, on=True):
        """Reuse the octree, lists and M2L classes while coordinates, ncrit and theta are unchanged."""
        check(self._lib.fmmh_set_option(self._h, 4, 1.0 if on else 0.0))

    def set_m2l_mfma(self, mode=1):
        """M2L engine: 0 per-target VALU, 1 auto (default: per-target MFMA over >= 16 sets, else the
        class-grouped MFMA GEMM), 2 per-target MFMA for any set count, 3 class-grouped always."""
        check(self._lib.fmmh_set_option(self._h, 3, float(mode)))

    def phase_times(self):
        ms = np.zeros(PHASE_CAP)
        names = ctypes.create_string_buffer(4096)
        cnt = self._lib.fmmh_phase_times(self._h, _p(ms), PHASE_CAP, names, 4096)
        if cnt < 0:
            check(-cnt)
        keys = names.value.decode().split(";") if names.value else []
        return dict(zip(keys, ms[:cnt].tolist()))

    # -- learned per-pair expansion orders --------------------------------------
    def train_selector(self, coordinates, multipoles, theta, ncrit, cap, nsamples=4096, epochs=40, seed=1):
        """Measure M2L truncation errors on a sample system and train the order
        selector on the GPU.  Returns a dict of training figures."""
        X = np.asfortranarray(coordinates, dtype=np.float64)
        M = np.asfortranarray(multipoles, dtype=np.float64)
        n, dim = M.shape
        info = np.zeros(8)
        check(self._lib.fmmh_selector_train(self._h, _p(X), n, _p(M), n, n, dim, float(theta), int(ncrit), int(cap),
                                            int(nsamples), int(epochs), int(seed), _p(info)))
        keys = ("examples", "samples", "first_loss", "final_loss", "sample_ms", "train_ms", "steps",
                "graph_replays")
        return dict(zip(keys, info.tolist()))

    def selector_samples(self, nsamples, cap):
        pairs = np.zeros((nsamples, 2), np.int32)
        feat = np.zeros((nsamples, 4), np.float32)
        err = np.zeros((nsamples, cap + 1))
        fref = np.zeros(nsamples)
        check(self._lib.fmmh_selector_samples(self._h, _p(pairs), _p(feat), _p(err), _p(fref)))
        return pairs, feat, err, fref

    def selector_loss(self):
        cnt = self._lib.fmmh_selector_loss(self._h, None, 0)
        if cnt < 0:
            check(-cnt)
        out = np.zeros(max(cnt, 1), np.float32)
        self._lib.fmmh_selector_loss(self._h, _p(out), cnt)
        return out[:cnt]

    def selector_weights(self):
        W = np.zeros(self._lib.fmmh_selector_nparam(), np.float32)
        norm = np.zeros(8, np.float32)
        meta = np.zeros(4)
        check(self._lib.fmmh_selector_get(self._h, _p(W), _p(norm), _p(meta)))
        return W, norm, meta

    def set_selector_weights(self, W, norm, meta):
        W = np.ascontiguousarray(W, np.float32)
        norm = np.ascontiguousarray(norm, np.float32)
        meta = np.ascontiguousarray(meta, np.float64)
        check(self._lib.fmmh_selector_set(self._h, _p(W), _p(norm), _p(meta)))

    def selector_grad(self, W, norm, X, Y, asym=4.0):
        """Mean loss and its gradient at W for the examples (X, Y), by the training kernels."""
        W = np.ascontiguousarray(W, np.float32)
        norm = np.ascontiguousarray(norm, np.float32)
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        out = np.zeros(W.size + 1, np.float32)
        check(self._lib.fmmh_selector_grad(self._h, _p(W), _p(norm), _p(X), _p(Y), X.shape[0], float(asym),
                                           _p(out)))
        return float(out[-1]), out[:-1]

    def selector_predict(self, W, norm, X):
        W = np.ascontiguousarray(W, np.float32)
        norm = np.ascontiguousarray(norm, np.float32)
        X = np.ascontiguousarray(X, np.float32)
        out = np.zeros(X.shape[0], np.float32)
        check(self._lib.fmmh_selector_predict(self._h, _p(W), _p(norm), _p(X), X.shape[0], _p(out)))
        return out

    def set_order_tolerance(self, tol, spot_checks=0, margin=None, safety=None, force_full=False):
        """tol > 0: learned per-pair M2L orders for a relative RMS far-field error of about tol
        (needs a trained selector); spot_checks > 0 validates each call and falls back to
        expansion_order when the check fails."""
        check(self._lib.fmmh_set_option(self._h, 6, float(tol)))
        check(self._lib.fmmh_set_option(self._h, 8, float(spot_checks)))
        if margin is not None:
            check(self._lib.fmmh_set_option(self._h, 7, float(margin)))
        if safety is not None:
            check(self._lib.fmmh_set_option(self._h, 9, float(safety)))
        check(self._lib.fmmh_set_option(self._h, 10, 1.0 if force_full else 0.0))

    def set_autotune(self, on=True, reset=False):
        """Tile-shape autotuner of the many-sets M2L (on by default); reset forgets earlier choices."""
        check(self._lib.fmmh_set_option(self._h, 11, -1.0 if reset else (1.0 if on else 0.0)))

    def set_near_chunk(self, cap=0):
        """Cap on the pair work of one near-field workgroup (source-tile iterations; 0: automatic)."""
        check(self._lib.fmmh_set_option(self._h, 13, float(cap)))

    def set_graph(self, on=True):
        """Replay tree-reuse calls of an unchanged shape as a captured HIP graph."""
        check(self._lib.fmmh_set_option(self._h, 12, 1.0 if on else 0.0))

    def tuning(self):
        """{key: variant} of the autotuner and the per-variant times (ms) of its last tuning run."""
        n = self._lib.fmmh_tuning(self._h, None, None, None, 0)
        if n < 0:
            check(-n)
        keys = np.zeros(max(n, 1), np.int64)
        var = np.zeros(max(n, 1), np.int32)
        ms = np.zeros(4)
        self._lib.fmmh_tuning(self._h, _p(keys), _p(var), _p(ms), n)
        return {int(k): int(v) for k, v in zip(keys[:n], var[:n])}, ms

    def selection_stats(self):
        out = np.zeros(21)
        check(self._lib.fmmh_selection_stats(self._h, _p(out), 21))
        return dict(spot_error=float(out[0]), fallback=bool(out[1]), tau=float(out[2]), m2l_flops=float(out[3]),
                    m2l_flops_full=float(out[4]), mean_order=float(out[5]), active=bool(out[6]),
                    near_rms=float(out[7]), order_hist=out[8:19].astype(np.int64), tolerance_scale=float(out[19]),
                    full_order_spot_error=float(out[20]))

    def stats(self):
        out = np.zeros(12, dtype=np.int64)
        check(self._lib.fmmh_stats(self._h, _p(out), 12))
        keys = ("num_nodes", "num_leaves", "max_depth", "n_m2l", "n_p2p", "p2p_site_pairs", "m2l_classes",
                "list_retries", "graph", "graph_replays", "near_workgroups", "host_waits")
        return dict(zip(keys, (int(v) for v in out)))

    def debug_buffer(self, name, count, dtype=np.float64):
        """First `count` elements of an internal device buffer (white-box tests)."""
        out = np.zeros(count, dtype=dtype)
        check(self._lib.fmmh_debug_buffer(self._h, name.encode(), _p(out), out.nbytes))
        return out

    def tree(self, coordinates, ncrit, theta):
        """Octree + interaction lists (pre-order ids, 0-based) as numpy arrays."""
        X = np.asfortranarray(coordinates, dtype=np.float64)
        n = X.shape[0]
        info = np.zeros(5, dtype=np.int64)
        check(self._lib.fmmh_tree_build(self._h, _p(X), n, n, int(ncrit), float(theta), _p(info)))
        nn, nl, md, nm2l, np2p = (int(v) for v in info)
        out = dict(centers=np.zeros((nn, 3), order="F"), r=np.zeros(nn), rmax=np.zeros(nn),
                   depth=np.zeros(nn, np.int32), parent=np.zeros(nn, np.int32), is_leaf=np.zeros(nn, np.int32),
                   nleaf=np.zeros(nn, np.int32), mem_off=np.zeros(nn + 1, np.int32),
                   members=np.zeros(n, np.int32), m2l_off=np.zeros(nn + 1, np.int32),
                   m2l_src=np.zeros(max(nm2l, 1), np.int32), p2p_off=np.zeros(nn + 1, np.int32),
                   p2p_src=np.zeros(max(np2p, 1), np.int32))
        check(self._lib.fmmh_tree_export(self._h, *[_p(out[k]) for k in (
            "centers", "r", "rmax", "depth", "parent", "is_leaf", "nleaf", "mem_off", "members", "m2l_off",
            "m2l_src", "p2p_off", "p2p_src")]))
        out["m2l_src"] = out["m2l_src"][:nm2l]
        out["p2p_src"] = out["p2p_src"][:np2p]
        out["max_depth"] = md
        out["num_leaves"] = nl
        return out


_default = None


def _ctx():
    global _default
    if _default is None:
        _default = Context()
    return _default


def field_direct(comm, coordinates, multipoles, exclusions, field_order, damp_type="", damping_factors=()):
    """Drop-in for octree.field_direct (src/fmm.F90:1338-1352)."""
    return _ctx().field_direct(comm, coordinates, multipoles, exclusions, field_order, damp_type, damping_factors)


def field_fmm(comm, coordinates, multipoles, exclusions, theta, ncrit, expansion_order, field_order, damp_type="",
              damping_factors=(), is_periodic=False, box_matrix=None):
    """Drop-in for octree.field_fmm (src/fmm.F90:1354-1400)."""
    return _ctx().field_fmm(comm, coordinates, multipoles, exclusions, theta, ncrit, expansion_order, field_order,
                            damp_type, damping_factors, is_periodic, box_matrix)


def get_multipoles(comm, coordinates, multipoles, target_coordinates, theta, ncrit, expansion_order):
    """Drop-in for fmm.get_multipoles (src/fmm.F90:1171-1229)."""
    return _ctx().get_multipoles(comm, coordinates, multipoles, target_coordinates, theta, ncrit, expansion_order)


def get_boxes(coordinates, target_coordinates, theta, ncrit, expansion_order):
    """Drop-in for fmm.get_boxes (src/fmm.F90:1231-1291)."""
    return _ctx().get_boxes(coordinates, target_coordinates, theta, ncrit, expansion_order)


def interpolate_electron_fields(comm, coordinates, local_expansion, box_indices, theta, ncrit, expansion_order,
                                field_order):
    """Drop-in for fmm.interpolate_electron_fields (src/fmm.F90:1293-1336)."""
    return _ctx().interpolate_electron_fields(comm, coordinates, local_expansion, box_indices, theta, ncrit,
                                              expansion_order, field_order)
