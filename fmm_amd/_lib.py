"""Loader for libfmm_hip.so.  There is no fallback: if the library is
missing the import of fmm_amd.octree fails loudly."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfmm_hip.so")

_c = ctypes
i32, i64, f64, vp = _c.c_int32, _c.c_int64, _c.c_double, _c.c_void_p

_PROTOS = {
    "fmmh_context_create": (i32, [_c.POINTER(vp), i32]),
    "fmmh_context_destroy": (i32, [vp]),
    "fmmh_last_error": (_c.c_char_p, []),
    "fmmh_parse_damp_type": (i32, [_c.c_char_p, i64, _c.POINTER(i32)]),
    "fmmh_field_direct": (i32, [vp, i32, vp, i64, vp, i64, i32, i32, vp, i64, i32, i32, vp, i64, i32, vp, i64]),
    "fmmh_field_direct_dev": (i32, [vp, i32, vp, i64, vp, i64, i32, i32, vp, i64, i32, i32, vp, i64, i32, vp, i64]),
    "fmmh_field_direct_dev_noalloc": (i32, [vp, i32, vp, i64, vp, i64, i32, i32, vp, i64, i32, i32, vp, i64, i32, vp,
                                            i64]),
    "fmmh_field_fmm": (i32, [vp, i32, vp, i64, vp, i64, i32, i32, vp, i64, i32, f64, i32, i32, i32, vp, i64, i32,
                             vp, i64, i32, vp]),
    "fmmh_field_fmm_dev": (i32, [vp, i32, vp, i64, vp, i64, i32, i32, vp, i64, i32, f64, i32, i32, i32, vp, i64,
                                 i32, vp, i64, i32, vp]),
    "fmmh_field_direct_sets": (i32, [vp, i32, vp, i64, vp, i64, i32, i32, i32, vp, i64, i32, i32, vp, i64, i32, vp,
                                     i64]),
    "fmmh_field_direct_sets_dev": (i32, [vp, i32, vp, i64, vp, i64, i32, i32, i32, vp, i64, i32, i32, vp, i64, i32,
                                         vp, i64]),
    "fmmh_field_fmm_sets": (i32, [vp, i32, vp, i64, vp, i64, i32, i32, i32, vp, i64, i32, f64, i32, i32, i32, vp, i64,
                                  i32, vp, i64, i32, vp]),
    "fmmh_field_fmm_sets_dev": (i32, [vp, i32, vp, i64, vp, i64, i32, i32, i32, vp, i64, i32, f64, i32, i32, i32, vp,
                                      i64, i32, vp, i64, i32, vp]),
    "fmmh_field_direct_targets_dev": (i32, [vp, vp, i64, vp, i64, i32, i32, vp, i64, i32, vp, i32, i32, vp, i64,
                                            i32, vp, i64]),
    "fmmh_get_boxes": (i32, [vp, vp, i64, i32, vp, i64, i32, f64, i32, i32, vp]),
    "fmmh_get_boxes_export": (i32, [vp, vp, vp, vp]),
    "fmmh_get_multipoles": (i32, [vp, i32, vp, i64, i32, vp, i64, i32, i32, vp, i64, i32, f64, i32, i32, vp]),
    "fmmh_get_multipoles_export": (i32, [vp, vp, vp, vp]),
    "fmmh_interpolate_electron_fields": (i32, [vp, i32, vp, i64, i32, vp, i64, i32, i32, vp, f64, i32, i32, i32, vp,
                                               i64]),
    "fmmh_tree_build": (i32, [vp, vp, i64, i32, i32, f64, vp]),
    "fmmh_tree_export": (i32, [vp] + [vp] * 13),
    "fmmh_set_option": (i32, [vp, i32, f64]),
    "fmmh_phase_times": (i32, [vp, vp, i32, _c.c_char_p, i32]),
    "fmmh_stats": (i32, [vp, vp, i32]),
    "fmmh_debug_buffer": (i32, [vp, _c.c_char_p, vp, i64]),
    "fmmh_selector_train": (i32, [vp, vp, i64, vp, i64, i32, i32, f64, i32, i32, i32, i32, _c.c_uint64, vp]),
    "fmmh_selector_samples": (i32, [vp, vp, vp, vp, vp]),
    "fmmh_selector_loss": (i32, [vp, vp, i32]),
    "fmmh_selector_nparam": (i32, []),
    "fmmh_selector_get": (i32, [vp, vp, vp, vp]),
    "fmmh_selector_set": (i32, [vp, vp, vp, vp]),
    "fmmh_selector_grad": (i32, [vp, vp, vp, vp, vp, i32, _c.c_float, vp]),
    "fmmh_selector_predict": (i32, [vp, vp, vp, vp, i32, vp]),
    "fmmh_selection_stats": (i32, [vp, vp, i32]),
    "fmmh_tuning": (i32, [vp, vp, vp, vp, i32]),
}

EXPORTED = tuple(_PROTOS)
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m fmm_amd.build` "
                              "(there is no CPU fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _PROTOS.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


class FmmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def check(rc):
    if rc != 0:
        raise FmmError(rc, load().fmmh_last_error().decode(errors="replace"))
    return rc
