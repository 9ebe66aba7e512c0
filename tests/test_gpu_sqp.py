"""minimize_SQP on the device (contact as the reference's hard constraints,
solver.cpp:109-197) against the oracle's SQP, whose QPs are solved by a
restatement of OSQP (the solver the reference calls) run to the exact QP
solution: the reference's own no-penetration scenario, a cube landing on the
ground plane, a d03-style disk poke with a clamped base (NeumannBC rows,
backward Euler and Crank-Nicolson), and reference-style drivers through the
py_sors drop-in (d03_pokeflex.py's disk, d05_hopping_leg.py's hex leg with
muscles on the ground).  Bound: 1e-6 of the bounding-box diagonal in mixed
precision, 1e-8 in FP64."""
import importlib
import os
import sys

import numpy as np
import pytest

from sors_amd import mesh, native

pytestmark = pytest.mark.gpu
PKG = os.path.join(os.path.dirname(os.path.dirname(__file__)), "sors_amd")
# both minimisers stop at ||dx|| < 1e-9 m (DX_TOL, common.h:25), which bounds
# how far two converged states can lie apart: ~1e-8 of these 5-10 cm boxes
PRECISIONS = [(native.FP64, 1e-8), (native.MIXED, 1e-6)]


def say(*a):
    print(*a, flush=True)


def bbox(X):
    X = np.asarray(X).reshape(-1, 3)
    return float(np.linalg.norm(X.max(0) - X.min(0)))


@pytest.mark.parametrize("precision,bound", PRECISIONS)
def test_reference_plane_contact_no_penetration(orc, precision, bound):
    """cpp/src/_test.cpp:629-698 on the device: after one SQP step no vertex
    is below z = 0 by more than TOL = 1e-4 (_test.h:210), and the step
    matches the oracle's."""
    X = np.array([[0, 0, -0.10], [1, 0, 0.20], [0, 1, 0.20], [0, 0, 0.30]], float)
    conn = np.array([[0, 1, 2, 3]], np.int32)
    say("plane contact known answer, precision", precision)
    r = native.Rollout(X, conn, [1e5], [0.3], [1.0], grav=(0, 0, -9.81), method=native.MINIMIZE_SQP,
                       precision=precision)
    r.set_contact(planes=[((0, 0, 0), (0, 0, 1))], surface=[0, 1, 2, 3])
    o = orc.System(X, conn, [1e5], [0.3], [1.0], grav=(0, 0, -9.81), method=orc.SQP)
    o.set_contact(planes=[((0, 0, 0), (0, 0, 1))], surface=[0, 1, 2, 3])
    q, st = r.step(X.reshape(-1).copy(), 1e-2)
    qo, _ = o.step(X.reshape(-1).copy(), 1e-2)
    err = np.abs(q - qo).max() / bbox(X)
    say("  newton", st.newton_iters, "qp", st.qp_iters, "pcg", st.pcg_iters, "err/bbox", err)
    assert q.reshape(-1, 3)[:, 2].min() >= -1e-4
    assert abs(q[2]) <= 1e-9          # the penetrating vertex sits on the plane (its row active)
    assert err <= bound


def cube_scene(n=4, h=0.02):
    X, hexes = mesh.box_grid(n, n, n, h, h, h)
    conn = mesh.kuhn_split(hexes)
    surf = mesh.box_surface_nodes(X)
    X[:, 2] += 1e-3
    return X, conn, surf


@pytest.mark.parametrize("precision,bound", PRECISIONS)
def test_cube_landing_on_plane_matches_oracle(orc, precision, bound):
    """A 4^3-cell tet cube 1 mm above the ground moving down at 0.3 m/s under
    gravity: it lands, is held by the plane rows and bounces; 25 SQP steps."""
    X, conn, surf = cube_scene()
    E = len(conn)
    args = (X, conn, np.full(E, 2e5), np.full(E, 0.4), np.full(E, 1000.0))
    v0 = np.zeros_like(X)
    v0[:, 2] = -0.3
    dt = 4e-3
    say("cube landing, precision", precision)
    r = native.Rollout(*args, grav=(0, 0, -9.81), method=native.MINIMIZE_SQP, precision=precision)
    r.set_contact(planes=[((0, 0, 0), (0, 0, 1))], surface=surf)
    r.set_initial_velocity(v0.reshape(-1))
    o = orc.System(*args, grav=(0, 0, -9.81), method=orc.SQP)
    o.set_contact(planes=[((0, 0, 0), (0, 0, 1))], surface=surf)
    o.set_initial_velocity(v0.reshape(-1))
    q = qo = X.reshape(-1).copy()
    worst, zmin, qp = 0.0, 1.0, 0
    for k in range(25):
        q, st = r.step(q, dt)
        qo, _ = o.step(qo, dt)
        worst = max(worst, np.abs(q - qo).max() / bbox(X))
        zmin = min(zmin, q[2::3].min())
        qp += st.qp_iters
    say("  max err/bbox", worst, "lowest z", zmin, "working-set iterations", qp)
    assert zmin >= -1e-9 * bbox(X)
    assert worst <= bound


@pytest.mark.parametrize("scheme", [native.BACKWARD_EULER, native.CRANK_NICOLSON])
@pytest.mark.parametrize("precision,bound", PRECISIONS)
def test_disk_poke_clamped_base_matches_oracle(orc, scheme, precision, bound):
    """python/d03_pokeflex.py's setting: the bottom face clamped (NeumannBC
    equality rows), a disk (normal -z) pressing into the top face and moved
    with set_disk before every step; its captured footprint freezes."""
    p = mesh.contact_poke(n=5, steps=16, gap=1e-3, speed=0.75, jitter=5e-4, jitter_every=8)
    sc = p.scene
    clamp = np.repeat(sc.X[:, 2] < 1e-9, 3).astype(np.int32)
    normal = (0.0, 0.0, -1.0)
    say("disk poke, scheme", scheme, "precision", precision)
    r = native.Rollout(sc.X, sc.conn, sc.youngs, sc.poisson, sc.density, clamp, np.zeros(3 * sc.n_nodes), sc.grav,
                       scheme=scheme, method=native.MINIMIZE_SQP, precision=precision)
    r.set_contact(disks=[(p.centers[0], normal, p.radius)], surface=p.surface)
    o = orc.System(sc.X, sc.conn, sc.youngs, sc.poisson, sc.density, clamp, np.zeros(3 * sc.n_nodes), sc.grav,
                   scheme=orc.CN if scheme == native.CRANK_NICOLSON else orc.BE, method=orc.SQP)
    o.set_contact(disks=[(p.centers[0], normal, p.radius)], surface=p.surface)
    q = qo = sc.X.reshape(-1).copy()
    worst = 0.0
    for k in range(sc.steps):
        r.set_disk(0, p.centers[k + 1], normal, p.radius)
        o.set_disk(0, p.centers[k + 1], normal, p.radius)
        q, st = r.step(q, sc.dt)
        qo, _ = o.step(qo, sc.dt)
        worst = max(worst, np.abs(q - qo).max() / sc.bbox_diag)
    x = q.reshape(-1, 3)
    c = p.centers[-1]
    under = (sc.X[:, 2] > sc.X[:, 2].max() - 1e-9) & (np.hypot(sc.X[:, 0] - c[0], sc.X[:, 1] - c[1]) < p.radius)
    say("  max err/bbox", worst, "footprint z", x[under, 2].max(), "disk", c[2])
    assert under.sum() >= 1
    assert x[under, 2].max() <= c[2] + 1e-9 * sc.bbox_diag    # the footprint stays below the disk plane
    assert x[under, 2].max() < sc.X[:, 2].max() - 3e-3         # indented
    assert np.abs(x[sc.X[:, 2] < 1e-9] - sc.X[sc.X[:, 2] < 1e-9]).max() <= 1e-12  # base held
    assert worst <= bound


def load_pysors():
    sys.path.insert(0, PKG)
    try:
        return importlib.import_module("py_sors")
    finally:
        sys.path.remove(PKG)


def test_pokeflex_style_driver_with_constraints(orc):
    """python/d03_pokeflex.py through py_sors with the default contact model:
    neumannBC + diskContact under minimize_SQP, crank_nicolson, the disk
    moved through get_disk_contact_constraint() before every step."""
    ps = load_pysors()
    p = mesh.contact_poke(n=5, steps=12, gap=1e-3, speed=0.75, jitter=0.0)
    sc = p.scene
    clamp = sc.X[:, 2] < 1e-9
    mask = np.repeat(clamp, 3).astype(float)
    settings = ps.SimulationSettings()
    settings.meshType = "Tetrahedron"
    settings.timeSteppingScheme = "crank_nicolson"
    settings.solverMethod = "minimize_SQP"
    settings.dt = sc.dt
    assert settings.contactModel == "constraints"
    energies = [set(["neohookean"]) for _ in range(sc.n_elems)]
    params = [ps.Params(["density", "youngsModulus", "poissonsRatio"], [1100.0, 2e5, 0.4])
              for _ in range(sc.n_elems)]
    normal = np.array([0.0, 0.0, -1.0])
    cparams = ps.Params(["neumannBCmask", "neumannBCvalue", "diskCenterVectorList", "diskNormalVectorList",
                         "diskRadiusList", "surfaceVerticesIdx"],
                        [mask, np.zeros(3 * sc.n_nodes), p.centers[0], normal, [p.radius], p.surface.astype(float)])
    energy = ps.Energy(sc.X, sc.conn, settings, sc.grav, 0.0, energies, params, ["neumannBC", "diskContact"],
                       cparams, [], ps.Params())
    disk = ps.get_disk_contact_constraint(energy, 1)
    solver = ps.Solver(settings)
    o = orc.System(sc.X, sc.conn, sc.youngs, sc.poisson, sc.density, np.repeat(clamp, 3).astype(np.int32),
                   np.zeros(3 * sc.n_nodes), sc.grav, scheme=orc.CN, method=orc.SQP)
    o.set_contact(disks=[(p.centers[0], normal, p.radius)], surface=p.surface)
    q = qo = sc.X.reshape(-1).copy()
    for k in range(sc.steps):
        disk.set_disk(p.centers[k + 1], normal, p.radius, diskIdx=0)
        o.set_disk(0, p.centers[k + 1], normal, p.radius)
        q = solver.step(q, energy)
        qo, _ = o.step(qo, sc.dt)
    err = np.abs(q - qo).max() / sc.bbox_diag
    say("pokeflex-style SQP driver vs oracle", err)
    assert err <= 1e-6
    c = p.centers[-1]
    under = (sc.X[:, 2] > sc.X[:, 2].max() - 1e-9) & (np.hypot(sc.X[:, 0] - c[0], sc.X[:, 1] - c[1]) < p.radius)
    assert q.reshape(-1, 3)[under, 2].max() < sc.X[:, 2].max() - 2e-3


def test_hopping_leg_style_driver_with_constraints(orc):
    """python/d05_hopping_leg.py's structure through py_sors with the default
    contact model: hex elements, a muscle band in two groups, the ground plane
    as planeContact rows under minimize_SQP, an actuation ramp that presses
    the leg into the ground."""
    ps = load_pysors()
    sc = mesh.beam(nx=3, ny=3, nz=10, lx=0.03, ly=0.03, lz=0.1, element="hex", E=5e5, nu=0.4, rho=1000.0,
                   tip_speed=0.0, steps=16)
    X = sc.X + np.array([0.0, 0.0, 1e-4])
    settings = ps.SimulationSettings()
    settings.meshType = "Hexahedron"
    settings.timeSteppingScheme = "backward_euler"
    settings.solverMethod = "minimize_SQP"
    settings.dt = 2e-3
    cen = X[sc.conn].mean(axis=1)
    energies, params, group = [], [], np.full(sc.n_elems, -1, np.int32)
    for e in range(sc.n_elems):
        pr = ps.Params(["density", "youngsModulus", "poissonsRatio"], [1000.0, 5e5, 0.4])
        es = set(["neohookean"])
        if 0.02 < cen[e, 2] < 0.08:
            es.add("softconMuscle")
            group[e] = int(cen[e, 0] >= 0.015)
            pr.set_param("softconStiffness", 5e5)
            pr.set_param("softconDirection", [0.0, 0.0, 1.0])
            pr.set_param("muscleGroup", float(group[e]))
        energies.append(es)
        params.append(pr)
    surf = mesh.box_surface_nodes(X)
    cparams = ps.Params(["positionVectorListPlanes", "normalVectorListPlanes", "surfaceVerticesIdx"],
                        [[0.0, 0.0, 0.0], [0.0, 0.0, 1.0], surf.astype(float)])
    energy = ps.Energy(X, sc.conn, settings, sc.grav, 0.0, energies, params, ["planeContact"], cparams,
                       ["vertexForce"], ps.Params())
    solver = ps.Solver(settings)
    o = orc.System(X, sc.conn, sc.youngs, sc.poisson, sc.density, grav=sc.grav, method=orc.SQP)
    o.set_contact(planes=[((0.0, 0.0, 0.0), (0.0, 0.0, 1.0))], surface=surf)
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
    say("hopping-leg-style SQP driver vs oracle", err, "lowest z", q[2::3].min())
    assert q[2::3].min() >= -1e-9
    assert err <= 1e-6
