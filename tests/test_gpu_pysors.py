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
    before every step.  With contactModel "penalty" the device runs its
    penalty formulation; compared with the oracle's restatement of the same
    formulation."""
    ps = load()
    p = mesh.contact_poke(n=5, steps=20, gap=-5e-4, speed=0.05)
    sc = p.scene
    clamp = sc.X[:, 2] < 1e-9                       # fixed bottom (d03_pokeflex.py:113-121)
    mask = np.repeat(clamp, 3).astype(float)
    settings = ps.SimulationSettings()
    settings.meshType = "Tetrahedron"
    settings.timeSteppingScheme = "backward_euler"
    settings.solverMethod = "minimize_SQP"
    settings.contactModel = "penalty"
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


def test_stable_neohookean_cantilever_driver(orc):
    """The structure of python/d02_cantilever.py:32-38, 60-108: a clamped
    tet beam with "stableneohookean" per element, Crank-Nicolson,
    minimize_newton, two substeps, through the py_sors drop-in."""
    ps = load()
    sc = mesh.beam(nx=10, ny=2, nz=2, E=2.6e5, nu=0.49, rho=1070.0, steps=15, dt=2e-3)
    settings = ps.SimulationSettings()
    settings.meshType = "Tetrahedron"
    settings.timeSteppingScheme = "crank_nicolson"
    settings.solverMethod = "minimize_newton"
    settings.dt = sc.dt
    settings.substeps = 2
    energies = [set(["stableneohookean"]) for _ in range(sc.n_elems)]
    params = [ps.Params(["density", "youngsModulus", "poissonsRatio"], [1070.0, 2.6e5, 0.49])
              for _ in range(sc.n_elems)]
    cparams = ps.Params(["neumannBCmask", "neumannBCvalue"], [sc.bcmask.astype(float), sc.bcvalue])
    energy = ps.Energy(sc.X, sc.conn, settings, sc.grav, 0.0, energies, params, ["neumannBC"], cparams, [],
                       ps.Params())
    energy.set_initial_velocity(sc.v0.reshape(-1, 3))
    solver = ps.Solver(settings)
    o = orc.System.from_scene(sc, scheme=orc.CN, method=orc.NEWTON)
    o.set_material(orc.MAT_SNH)
    q = qo = sc.X.reshape(-1).copy()
    for _ in range(sc.steps):
        q = solver.step(q, energy)
        qo, _ = o.step(qo, sc.dt, substeps=2)
    err = np.abs(q - qo).max() / sc.bbox_diag
    print("stableneohookean cantilever driver vs oracle", err, flush=True)
    assert err <= 1e-6


def test_hopping_leg_style_muscle_driver(orc):
    """The structure of python/d05_hopping_leg.py:120-219, 331-360: hex
    elements with {"neohookean", "softconMuscle"} in a band, two muscle groups
    along +z ("muscleGroup", "softconStiffness", "softconDirection"), a ground
    plane ("planeContact" over the surface vertices), minimize_SQP, and a
    two-channel "softconMuscle" actuation ramp.  With contactModel "penalty"
    the device runs its penalty contact with plain Newton (the muscle energy
    is not the potential of its gradient); compared with the oracle's
    restatement of the same model."""
    ps = load()
    sc = mesh.beam(nx=3, ny=3, nz=10, lx=0.03, ly=0.03, lz=0.1, element="hex", E=5e5, nu=0.4, rho=1000.0,
                   tip_speed=0.0, steps=20)
    X = sc.X + np.array([0.0, 0.0, 1e-4])            # resting just above the ground plane z = 0
    settings = ps.SimulationSettings()
    settings.meshType = "Hexahedron"
    settings.timeSteppingScheme = "backward_euler"
    settings.solverMethod = "minimize_SQP"
    settings.contactModel = "penalty"
    settings.dt = 2e-3
    cen = X[sc.conn].mean(axis=1)
    energies, params, group = [], [], np.full(sc.n_elems, -1, np.int32)
    for e in range(sc.n_elems):
        p = ps.Params(["density", "youngsModulus", "poissonsRatio"], [1000.0, 5e5, 0.4])
        es = set(["neohookean"])
        if 0.02 < cen[e, 2] < 0.08:
            es.add("softconMuscle")
            group[e] = int(cen[e, 0] >= 0.015)
            p.set_param("softconStiffness", 5e5)
            p.set_param("softconDirection", [0.0, 0.0, 1.0])
            p.set_param("muscleGroup", float(group[e]))
        energies.append(es)
        params.append(p)
    surf = mesh.box_surface_nodes(X)
    cparams = ps.Params(["positionVectorListPlanes", "normalVectorListPlanes", "surfaceVerticesIdx",
                         "contactStiffness"], [[0.0, 0.0, 0.0], [0.0, 0.0, 1.0], surf.astype(float), [1e5]])
    energy = ps.Energy(X, sc.conn, settings, sc.grav, 0.0, energies, params, ["planeContact"], cparams,
                       ["vertexForce"], ps.Params())
    solver = ps.Solver(settings)
    o = orc.System(X, sc.conn, sc.youngs, sc.poisson, sc.density, grav=sc.grav, method=orc.NEWTON)
    o.set_contact(planes=[((0.0, 0.0, 0.0), (0.0, 0.0, 1.0))], surface=surf, stiffness=1e5)
    o.set_muscle(group, np.full(sc.n_elems, 5e5), np.tile([0.0, 0.0, 1.0], (sc.n_elems, 1)))
    q = qo = X.reshape(-1).copy()
    act = ps.Params(["softconMuscle", "vertexForce"], [np.zeros((2, 1)), np.zeros((sc.n_nodes, 3))])
    for k in range(sc.steps):
        a = np.array([[0.4 * min(1.0, k / 8.0)], [0.15 * min(1.0, k / 8.0)]])
        act.set_param("softconMuscle", a)
        o.set_muscle_activation(a.reshape(-1))
        q = solver.step(q, energy, act, 0.0)
        qo, _ = o.step(qo, settings.dt)
    err = np.abs(q - qo).max() / sc.bbox_diag
    moved = np.abs(qo - X.reshape(-1)).max()
    print("hopping-leg-style muscle driver vs oracle", err, "max displacement", moved, flush=True)
    assert moved > 1e-4
    assert err <= 1e-6


def test_run_msd_timestepping(orc, tmp_path):
    """py_sors.run_msd_timestepping (cpp/py_sors.cpp:27,
    cpp/demos/03_msd_timestepping.cpp:16-212) on the reference's
    msd_cylinder_v3.msh (a data file copied from cpp/demos/meshes): 2 steps;
    the VTU per step, the PVD and msd_displacement.csv are written, and the
    tracked centre-of-mass heights match an oracle rollout of the same
    system (Crank-Nicolson, minimize_newton, CFL substeps, mass damping 1)."""
    import csv

    ps = load()
    path = os.path.join(os.path.dirname(__file__), "golden", "meshes", "msd_cylinder_v3.msh")
    out = str(tmp_path)
    steps = 3
    ps.run_msd_timestepping(path, out, steps)
    names = sorted(os.listdir(out))
    for i in range(steps):
        assert f"msd_cylinder_v3.msh_{i}.vtu" in names
    assert any(n.endswith(".pvd") for n in names)
    rows = list(csv.reader(open(os.path.join(out, "msd_displacement.csv"))))
    assert [int(r[0]) for r in rows] == list(range(steps))
    com = np.array([float(r[2]) for r in rows])
    # the same system on the oracle
    V, T = ps.IO.read_tet_MSH(path)
    V = np.asarray(V) * 1e-3
    T = np.asarray(T, np.int32)
    N, E = len(V), len(T)
    lo, hi = V.min(0), V.max(0)
    mask = np.repeat(V[:, 2] > hi[2] - 1e-4, 3).astype(np.int32)
    o = orc.System(V, T, np.full(E, 12.5 * 263824.0), np.full(E, 0.499), np.full(E, 1070.0), mask, np.zeros(3 * N),
                   (0.0, 0.0, -9.81), 1.0, scheme=orc.CN, method=orc.NEWTON, nthreads=8)
    t1 = [i for i in range(N) if np.all(np.abs(V[i] - lo) < 1e-3)][-1]
    t2 = [i for i in range(N) if np.all(np.abs(V[i] - [hi[0], hi[1], hi[2] / 2]) < 1e-3)][-1]
    q = V.reshape(-1).copy()
    ref = [(V[t1, 2] + V[t2, 2]) / 2]
    for _ in range(1, steps):
        dtc = o.cfl_timestep(q, 1e-2)
        q, _ = o.step(q, 1e-2, int(np.ceil(1e-2 / dtc)), linear_solver=orc.PCG, project=1, pcg_rtol=1e-10,
                      nthreads=8)
        ref.append((q[3 * t1 + 2] + q[3 * t2 + 2]) / 2)
    err = np.abs(com - np.array(ref)).max() / np.linalg.norm(hi - lo)
    print("msd demo: com heights", com, "vs oracle", err, flush=True)
    assert com[-1] < com[0]          # sags under gravity
    assert err <= 1e-6
