"""The C-ABI library loads and exports every entry point declared in
include/sors_amd.h (no device calls: runs without a GPU)."""
import ctypes
import os

from sors_amd import native


def test_header_declares_entry_points():
    syms = native.declared_symbols()
    for s in ("sors_create", "sors_destroy", "sors_step", "sors_step_resident", "sors_get_scatter_map",
              "sors_element_eval", "sors_linear_solve", "sors_last_error"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    assert os.path.exists(native.LIB_PATH), "run python -m sors_amd.build"
    lib = ctypes.CDLL(native.LIB_PATH)
    missing = [s for s in native.declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_default_options_match_reference_constants():
    o = native.default_options()
    # MAX_ITER = 100, DX_TOL = 1e-9 (cpp/headers/common.h:21-25)
    assert o.max_newton == 100 and o.dx_tol == 1e-9
    assert o.project_spd == 1 and o.precision == native.MIXED
