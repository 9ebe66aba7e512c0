"""The py_sors drop-in module: names, defaults and host-side behaviour of
the reference's Python surface (cpp/py_sors.cpp:25-261).  No device calls."""
import importlib
import os
import sys

import numpy as np
import pytest

PKG = os.path.join(os.path.dirname(os.path.dirname(__file__)), "sors_amd")


@pytest.fixture(scope="module")
def ps():
    sys.path.insert(0, PKG)
    try:
        return importlib.import_module("py_sors")
    finally:
        sys.path.remove(PKG)


def test_names(ps):
    for name in ("Energy", "EnergyTet", "EnergyHex", "Solver", "SolverTet", "SolverHex", "Params",
                 "SimulationSettings", "IO"):
        assert hasattr(ps, name)


def test_settings_defaults_match_reference(ps):
    # cpp/headers/simulationSettings.h:18-38
    s = ps.SimulationSettings()
    assert np.isnan(s.dt) and s.numSteps == -1 and s.substeps == 1
    assert s.timeSteppingScheme == "crank_nicolson" and s.solverMethod == "minimize_newton"
    assert s.CFLtimeSteppingFlag is False and s.numThreads == 1 and s.verbose == 1


def test_params(ps, tmp_path):
    p = ps.Params(["density", "youngsModulus"], [1070.0, 2e5])
    assert p.get_value("density")[0, 0] == 1070.0
    p.set_param("mask", np.array([1.0, 0.0, 1.0]))
    assert p.get_value("mask").shape == (3, 1)
    with pytest.raises(Exception):
        p.set_param("mask", np.zeros(4))  # an existing key keeps its size (params.h:43-49)
    f = tmp_path / "p.csv"
    f.write_text("density,1100\nsoftconDirection,0,0,1\n")
    q = ps.Params()
    q.set_param(str(f))
    assert q.get_value("density")[0, 0] == 1100.0
    assert q.get_value("softconDirection").ravel().tolist() == [0.0, 0.0, 1.0]
    with pytest.raises(Exception):
        q.get_value("missing")


def test_io_statics(ps):
    V, H = ps.IO.make_voxel_hex_grid(3, 2, 1, 0.1, 0.2, 0.3)
    assert V.shape == (24, 3) and H.shape == (6, 8)
    # io.cpp:946-947: vid = i + NX*(j + NY*k); corner 6 of cell 0 is (1,1,1)
    assert np.allclose(V[H[0, 6]], [0.1, 0.2, 0.3])
    T = ps.IO.hex_to_5_tets(H)
    assert T.shape == (30, 4) and T[0].tolist() == [H[0, 0], H[0, 1], H[0, 3], H[0, 4]]


def test_unknown_mesh_type_raises(ps):
    s = ps.SimulationSettings()
    s.meshType = "Prism"
    with pytest.raises(RuntimeError, match="Unknown mesh type"):
        ps.Solver(s)
