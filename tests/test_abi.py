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


def test_default_options_fill_every_field():
    """sors_default_options writes the struct's last fields where the ctypes
    mirror has them (sentinels overwritten): the two layouts agree, and the
    forcing-term extensions are off by default."""
    o = native.Options()
    o.coarse_nodes, o.coarse_every, o.pcg_rtol_first, o.reuse_tangent = -7, -7, 7.0, 7
    native.lib().sors_default_options(ctypes.byref(o))
    assert o.coarse_nodes == 2048 and o.coarse_every == 16
    assert o.pcg_rtol_first == 0.0 and o.reuse_tangent == 0


def test_server_declarations_and_option_checks():
    syms = native.declared_symbols()
    for s in ("sors_server_create", "sors_server_submit", "sors_server_wait", "sors_server_get_stats",
              "sors_server_destroy", "sors_server_default_options", "sors_server_last_error"):
        assert s in syms
    L = native.lib()
    o = native.ServerOptions()
    L.sors_server_default_options(ctypes.byref(o))
    assert o.max_active == 64 and o.hbm_budget_bytes == 200e9 and o.pcg_chunk % 2 == 0
    assert o.solver.max_newton == 100
    # option validation happens before any device call
    h = ctypes.c_void_p()
    o.max_active = 0
    assert L.sors_server_create(ctypes.byref(o), ctypes.byref(h)) == native.ERR_ARG
    assert b"max_active" in L.sors_server_last_error(None)
    o.max_active, o.pcg_chunk = 4, 1
    assert L.sors_server_create(ctypes.byref(o), ctypes.byref(h)) == native.ERR_ARG


def test_struct_sizes_match_header():
    # layouts the ctypes mirror relies on (x86-64: 8-byte aligned doubles)
    assert ctypes.sizeof(native.RequestStats) == 4 + 4 + 3 * 8 + 4 * 8 + 8  # + fp64_fallbacks and padding
    assert ctypes.sizeof(native.ServerStats) == 3 * 8 + 4 + 4 + 2 * 8 + 2 * 8
    assert ctypes.sizeof(native.ServerOptions) == 8 + 3 * 4 + 4 + ctypes.sizeof(native.Options)


def test_contact_entry_points_reject_null_context():
    L = native.lib()
    assert L.sors_set_contact(None, None) == native.ERR_ARG
    assert L.sors_set_disk(None, 0, None, None, 0.0) == native.ERR_ARG
    # ContactDesc mirrors sors_contact_desc: 3 counts, 6 array pointers, 3 doubles
    assert ctypes.sizeof(native.ContactDesc) == 3 * 8 + 6 * 8 + 3 * 8
