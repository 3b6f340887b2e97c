"""The C ABI library loads and exports every entry point that
include/fmm_hip.h declares (no compute calls: runs without a GPU)."""
import os
import re

import pytest

from conftest import ROOT


def _declared():
    src = open(os.path.join(ROOT, "include", "fmm_hip.h")).read()
    return sorted(set(re.findall(r"\b(fmmh_[a-z_0-9]+)\s*\(", src)))


def test_header_declares_the_drop_in_entry_points():
    names = _declared()
    for n in ("fmmh_field_direct", "fmmh_field_fmm", "fmmh_last_error", "fmmh_parse_damp_type"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from fmm_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libfmm_hip.so not built: run `python -m fmm_amd.build`")
    lib = _lib.load()
    for n in _declared():
        assert hasattr(lib, n), n
    assert set(_declared()) == set(_lib.EXPORTED)


def test_damp_type_parsing_matches_reference_rules():
    """trim(damp_type) compared case-sensitively (src/fmm.F90:585-595, 1372-1377)."""
    from fmm_amd.octree import FmmError, _damp_code
    assert _damp_code("") == 0
    assert _damp_code("      ") == 0
    assert _damp_code("THOLE ") == 1
    assert _damp_code("AMOEBA") == 2
    assert _damp_code("ERF   ") == 3
    for bad in ("thole", " THOLE", "TT", "ERF2"):
        with pytest.raises(FmmError, match="Unknown damp type"):
            _damp_code(bad)
