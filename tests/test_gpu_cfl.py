"""CFL sub-stepping (Solver::step with settings.CFLtimeSteppingFlag,
solver.cpp:16-20, and calculate_CFL_timestep, solver.cpp:300-331) on the
device, against the oracle."""
import importlib
import math
import os
import sys

import numpy as np
import pytest

from sors_amd import mesh, native

pytestmark = pytest.mark.gpu
PKG = os.path.join(os.path.dirname(os.path.dirname(__file__)), "sors_amd")


def say(*a):
    print(*a, flush=True)


def load_pysors():
    sys.path.insert(0, PKG)
    try:
        return importlib.import_module("py_sors")
    finally:
        sys.path.remove(PKG)


@pytest.mark.parametrize("element", ["tet", "hex"])
def test_device_cfl_matches_oracle(orc, element):
    sc = mesh.beam(nx=8, ny=3, nz=3, element=element, e_range=(5e5, 5e6), seed=2)
    rng = np.random.default_rng(4)
    v = rng.normal(scale=0.3, size=3 * sc.n_nodes)
    q = sc.X.reshape(-1) + rng.normal(scale=2e-4, size=3 * sc.n_nodes)
    r = native.Rollout.from_scene(sc)
    r.set_initial_velocity(v)
    for dt in (1e-3, 1.0):
        got = r.cfl_timestep(q, dt)
        ref = orc.cfl_timestep(q, sc.conn, v, dt)
        say(element, dt, got, ref)
        assert got == pytest.approx(ref, rel=1e-14)
        # the resident positions (the scene's rest shape before any step)
        assert r.cfl_timestep(None, dt) == pytest.approx(orc.cfl_timestep(sc.X, sc.conn, v, dt), rel=1e-14)
    assert r.cfl_timestep(q, 1e-9) == 1e-9  # capped at dt
    r.close()


def test_device_cfl_every_vertex_pair():
    # bottom face A(0,0) B(1,0) C(1.9,0.2) D(0.9,0.2), extruded by 1: the face
    # diagonal BD = sqrt(0.05) is shorter than every edge
    f = np.array([[0, 0], [1, 0], [1.9, 0.2], [0.9, 0.2]])
    X = np.zeros((8, 3))
    X[:4, :2] = f
    X[4:, :2] = f
    X[4:, 2] = 1.0
    r = native.Rollout(X, np.arange(8, dtype=np.int32)[None], [1e5], [0.3], [1000.0])
    v = np.zeros(24)
    v[7] = -2.0
    r.set_initial_velocity(v)
    assert r.cfl_timestep(X.reshape(-1), 1.0) == pytest.approx(np.sqrt(0.05) / 2.0, rel=1e-15)
    r.set_initial_velocity(np.zeros(24))
    assert r.cfl_timestep(X.reshape(-1), 0.5) == 0.5  # no motion: dt
    r.close()


def cfl_driver(ps, sc, v0, cfl, substeps=1):
    settings = ps.SimulationSettings()
    settings.meshType = "Tetrahedron"
    settings.timeSteppingScheme = "backward_euler"
    settings.solverMethod = "minimize_newton_linesearch"
    settings.dt = sc.dt
    settings.substeps = substeps
    settings.CFLtimeSteppingFlag = cfl
    energies = [set(["neohookean"]) for _ in range(sc.n_elems)]
    params = [ps.Params(["density", "youngsModulus", "poissonsRatio"], [1100.0, 1e6, 0.45])
              for _ in range(sc.n_elems)]
    cparams = ps.Params(["neumannBCmask", "neumannBCvalue"], [sc.bcmask.astype(float), sc.bcvalue])
    energy = ps.Energy(sc.X, sc.conn, settings, sc.grav, 0.0, energies, params, ["neumannBC"], cparams, [],
                       ps.Params())
    energy.set_initial_velocity(v0.reshape(-1, 3))
    return energy, ps.Solver(settings)


def test_pysors_cfl_substeps(orc):
    """With the flag on, a step takes ceil(dt / dt_CFL) substeps: the same
    state, bit for bit, as a step told that many substeps; several steps
    track the oracle stepped with its own CFL substep count."""
    ps = load_pysors()
    sc = mesh.beam(nx=10, ny=2, nz=2, element="tet", steps=4)
    # velocity scaled so that dt_CFL is about dt / 3.4
    h = orc.cfl_timestep(sc.X, sc.conn, np.ones(3 * sc.n_nodes), np.inf)
    x = sc.X[:, 0] / sc.X[:, 0].max()
    v0 = np.zeros((sc.n_nodes, 3))
    v0[:, 2] = 3.4 * h / sc.dt * x
    n0 = math.ceil(sc.dt / orc.cfl_timestep(sc.X, sc.conn, v0, sc.dt))
    assert n0 >= 3
    e1, s1 = cfl_driver(ps, sc, v0, True)
    e2, s2 = cfl_driver(ps, sc, v0, False, substeps=n0)
    q0 = sc.X.reshape(-1).copy()
    q1 = s1.step(q0, e1, 0.0)
    q2 = s2.step(q0, e2, 0.0)
    assert np.array_equal(q1, q2)
    o = orc.System.from_scene(sc, method=orc.NEWTON_LS)
    o.set_initial_velocity(v0.reshape(-1))
    q, qo = q0, q0
    counts = []
    for _ in range(sc.steps):
        n = math.ceil(sc.dt / o.cfl_timestep(qo, sc.dt))
        counts.append(n)
        qo, _ = o.step(qo, sc.dt, substeps=n)
    e3, s3 = cfl_driver(ps, sc, v0, True)
    for _ in range(sc.steps):
        q = s3.step(q, e3, 0.0)
    err = np.abs(q - qo).max() / sc.bbox_diag
    say("cfl substeps per step (oracle)", counts, "device vs oracle", err)
    assert err <= 1e-6
