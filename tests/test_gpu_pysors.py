"""A reference-style Python driver (the structure of python/d02_cantilever.py)
running on the device through the py_sors drop-in, compared with the FP64
oracle."""
import importlib
import os
import sys

import numpy as np
import pytest

from sors_amd import mesh

pytestmark = pytest.mark.gpu
PKG = os.path.join(os.path.dirname(os.path.dirname(__file__)), "sors_amd")


def load():
    sys.path.insert(0, PKG)
    try:
        return importlib.import_module("py_sors")
    finally:
        sys.path.remove(PKG)


@pytest.mark.parametrize("mesh_type", ["Tetrahedron", "Hexahedron"])
def test_cantilever_driver(orc, mesh_type):
    ps = load()
    element = "tet" if mesh_type == "Tetrahedron" else "hex"
    sc = mesh.beam(nx=10, ny=2, nz=2, element=element, steps=20)
    settings = ps.SimulationSettings()
    settings.meshType = mesh_type
    settings.timeSteppingScheme = "backward_euler"
    settings.solverMethod = "minimize_newton_linesearch"
    settings.dt = sc.dt
    energies = [set(["neohookean"]) for _ in range(sc.n_elems)]
    params = [ps.Params(["density", "youngsModulus", "poissonsRatio"], [1100.0, 1e6, 0.45])
              for _ in range(sc.n_elems)]
    cparams = ps.Params(["neumannBCmask", "neumannBCvalue"], [sc.bcmask.astype(float), sc.bcvalue])
    energy = ps.Energy(sc.X, sc.conn, settings, sc.grav, 0.0, energies, params, ["neumannBC"], cparams,
                       ["vertexForce"], ps.Params())
    energy.set_initial_velocity(sc.v0.reshape(-1, 3))
    solver = ps.Solver(settings)
    o = orc.System.from_scene(sc, method=orc.NEWTON_LS)
    q = qo = sc.X.reshape(-1).copy()
    act = ps.Params(["vertexForce"], [np.zeros((sc.n_nodes, 3))])
    for _ in range(sc.steps):
        q = solver.step(q, energy, act, 0.0)   # dt = 0 -> settings.dt (solver.cpp:8-9)
        qo, _ = o.step(qo, sc.dt)
    assert np.abs(q - qo).max() <= 1e-6 * sc.bbox_diag
    assert solver.last_stats["newton_iters"] >= 1


def test_inverted_element_message(orc):
    ps = load()
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) * 0.01
    settings = ps.SimulationSettings()
    settings.meshType = "Tetrahedron"
    settings.timeSteppingScheme = "backward_euler"
    settings.dt = 1e-3
    energy = ps.Energy(X, np.array([[0, 1, 2, 3]]), settings, np.zeros(3), 0.0, [set(["neohookean"])],
                       [ps.Params(["density", "youngsModulus", "poissonsRatio"], [1000.0, 1e5, 0.3])], [],
                       ps.Params(), [], ps.Params())
    solver = ps.Solver(settings)
    q = X.copy()
    q[3] = [0, 0, -0.01]  # node 3 pushed through the opposite face
    with pytest.raises(RuntimeError, match="Inverted Elements"):
        solver.step(q.reshape(-1), energy, 1e-3)
