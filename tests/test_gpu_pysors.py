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


@pytest.mark.parametrize("mesh_type,coarse", [("Tetrahedron", 0), ("Hexahedron", 0), ("Tetrahedron", 16),
                                              ("Hexahedron", 16)])
def test_cantilever_driver(orc, mesh_type, coarse):
    """coarse > 0: settings.coarseNodes turns on the two-level PCG."""
    ps = load()
    element = "tet" if mesh_type == "Tetrahedron" else "hex"
    sc = mesh.beam(nx=10, ny=2, nz=2, element=element, steps=20)
    settings = ps.SimulationSettings()
    settings.meshType = mesh_type
    settings.timeSteppingScheme = "backward_euler"
    settings.solverMethod = "minimize_newton_linesearch"
    settings.dt = sc.dt
    settings.coarseNodes = coarse
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
    pcg = 0
    for _ in range(sc.steps):
        q = solver.step(q, energy, act, 0.0)   # dt = 0 -> settings.dt (solver.cpp:8-9)
        qo, _ = o.step(qo, sc.dt)
        pcg += solver.last_stats["pcg_iters"]
    print(mesh_type, "coarse", coarse, "pcg iterations", pcg, flush=True)
    assert np.abs(q - qo).max() <= 1e-6 * sc.bbox_diag
    assert solver.last_stats["newton_iters"] >= 1


def test_frozen_hessian_driver(orc):
    """settings.solverMethod = "minimize_frozenhessian" through the drop-in,
    against the oracle's frozen iteration."""
    ps = load()
    sc = mesh.beam(nx=10, ny=2, nz=2, element="tet", steps=10)
    settings = ps.SimulationSettings()
    settings.meshType = "Tetrahedron"
    settings.timeSteppingScheme = "backward_euler"
    settings.solverMethod = "minimize_frozenhessian"
    settings.dt = sc.dt
    energies = [set(["neohookean"]) for _ in range(sc.n_elems)]
    params = [ps.Params(["density", "youngsModulus", "poissonsRatio"], [1100.0, 1e6, 0.45])
              for _ in range(sc.n_elems)]
    cparams = ps.Params(["neumannBCmask", "neumannBCvalue"], [sc.bcmask.astype(float), sc.bcvalue])
    energy = ps.Energy(sc.X, sc.conn, settings, sc.grav, 0.0, energies, params, ["neumannBC"], cparams, [],
                       ps.Params())
    energy.set_initial_velocity(sc.v0.reshape(-1, 3))
    solver = ps.Solver(settings)
    o = orc.System.from_scene(sc, method=orc.FROZEN)
    q = qo = sc.X.reshape(-1).copy()
    for _ in range(sc.steps):
        q = solver.step(q, energy, 0.0)
        qo, _ = o.step(qo, sc.dt)
    assert np.abs(q - qo).max() <= 1e-6 * sc.bbox_diag
    assert solver.last_stats["newton_iters"] >= 2


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


def test_pokeflex_style_disk_contact_driver(orc):
    """The structure of python/d03_pokeflex.py: neumannBC + diskContact, the
    SQP solver setting, the disk moved through get_disk_contact_constraint()
    before every step.  The device runs its penalty formulation; compared
    with the oracle's restatement of the same formulation."""
    ps = load()
    p = mesh.contact_poke(n=5, steps=20, gap=-5e-4, speed=0.05)
    sc = p.scene
    clamp = sc.X[:, 2] < 1e-9                       # fixed bottom (d03_pokeflex.py:113-121)
    mask = np.repeat(clamp, 3).astype(float)
    settings = ps.SimulationSettings()
    settings.meshType = "Tetrahedron"
    settings.timeSteppingScheme = "backward_euler"
    settings.solverMethod = "minimize_SQP"
    settings.dt = sc.dt
    energies = [set(["neohookean"]) for _ in range(sc.n_elems)]
    params = [ps.Params(["density", "youngsModulus", "poissonsRatio"], [1100.0, 2e5, 0.4])
              for _ in range(sc.n_elems)]
    normal = np.array([0.0, 0.0, -1.0])
    cparams = ps.Params(["neumannBCmask", "neumannBCvalue", "diskCenterVectorList", "diskNormalVectorList",
                         "diskRadiusList", "surfaceVerticesIdx", "contactStiffness", "frictionCoefficient"],
                        [mask, np.zeros(3 * sc.n_nodes), p.centers[0], normal, [p.radius], p.surface.astype(float),
                         [p.stiffness], [p.friction]])
    energy = ps.Energy(sc.X, sc.conn, settings, sc.grav, 0.0, energies, params, ["neumannBC", "diskContact"],
                       cparams, [], ps.Params())
    disk = ps.get_disk_contact_constraint(energy, 1)
    with pytest.raises(ValueError):
        ps.get_disk_contact_constraint(energy, 0)
    solver = ps.Solver(settings)
    o = orc.System(sc.X, sc.conn, sc.youngs, sc.poisson, sc.density, np.repeat(clamp, 3).astype(np.int32),
                   np.zeros(3 * sc.n_nodes), sc.grav, method=orc.NEWTON_LS)
    o.set_contact(disks=[(p.centers[0], normal, p.radius)], surface=p.surface, stiffness=p.stiffness,
                  friction=p.friction, eps_v=p.eps_v)
    q = qo = sc.X.reshape(-1).copy()
    for k in range(sc.steps):
        disk.set_disk(p.centers[k + 1], normal, p.radius, diskIdx=0)
        o.set_disk(0, p.centers[k + 1], normal, p.radius)
        q = solver.step(q, energy)
        qo, _ = o.step(qo, sc.dt)
    err = np.abs(q - qo).max() / sc.bbox_diag
    print("pokeflex-style driver vs oracle", err, flush=True)
    assert err <= 1e-6
    top = q.reshape(-1, 3)[:, 2].max()
    assert top < sc.X[:, 2].max()  # indented
