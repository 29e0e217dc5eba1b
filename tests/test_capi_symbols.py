"""The C-ABI library loads without a GPU and exports every entry point that
include/meshode.h declares (no compute calls here)."""
import ctypes as C
import os
import re
import subprocess

from meshode_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    src = open(os.path.join(ROOT, "include", "meshode.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mo_[a-z0-9_]+)\s*\(", src)))


def test_header_matches_loader_table():
    assert _header_symbols() == sorted(_lib.SYMBOLS)


def test_library_exports_all_symbols():
    lib = _lib.load()
    for name in _header_symbols():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (mo_[a-z0-9_]+)", out))
    assert set(_header_symbols()) <= exported
    # nothing else leaks out of the C ABI
    assert exported == set(_header_symbols())


def test_host_only_calls():
    lib = _lib.load()
    assert lib.mo_version() >= 100
    o = _lib.default_lm_options()
    assert o.max_num_iterations == 100 and o.initial_trust_region_radius == 1e4
    assert o.function_tolerance == 1e-6 and o.min_relative_decrease == 1e-3
    # an unknown template id is rejected before touching the device
    N = C.c_int32()
    rc = lib.mo_template_info(12345, C.byref(N), None, None)
    assert rc == -1 and b"param_id" in lib.mo_last_error()


def test_lm_shape_table_host_calls():
    """The LM launch-shape table is host state: shapes are listed, a fitting
    ascending table installs, and bad tables are rejected (no device work)."""
    import pytest
    from meshode_amd import autotune
    shp = autotune.shapes()
    assert len(shp) == 13 and shp[0] == (1024, 1) and max(t * v for t, v in shp) == autotune.LDS_MAX_VERTICES
    assert autotune.size_classes() == [1024, 2048, 4096, 4608, 5120, 5376, 6144, 8192]
    try:
        autotune.set_table([(1024, 4), (5120, 3)])          # 512x8 for <= 1024, 1024x5 for <= 5120
        with pytest.raises(_lib.MoError):
            autotune.set_table([(2048, 0)])                 # 1024x1 holds 1024 vertices only
        with pytest.raises(_lib.MoError):
            autotune.set_table([(4096, 2), (2048, 1)])      # not ascending
        with pytest.raises(_lib.MoError):
            autotune.set_table([(100, 13)])                 # no such shape
        with pytest.raises(_lib.MoError):
            _lib.call("mo_lm_shape_info", 13, None, None)
    finally:
        autotune.set_table([])
