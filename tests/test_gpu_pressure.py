"""PressureForce on the device (cpp/src/forces/pressureForce.cpp:14-185)
against the FP64 oracle: the gradient (-f(q)), the energy (+f(q).(q -
qPrev)) and the exported Hessian with the reference's force Jacobian under
minimize_newton (p/6 blocks, quads included); trajectories of the reference's
pressure test cube (cpp/src/_test.cpp:310-382) and of a SoPrA-style driver
(python/d04_sopra.py:73-113, 186-190) through py_sors."""
import importlib
import os
import sys

import numpy as np
import pytest

from sors_amd import mesh, native

pytestmark = pytest.mark.gpu
PKG = os.path.join(os.path.dirname(os.path.dirname(__file__)), "sors_amd")

HEX_FACES = [[0, 3, 2, 1], [4, 5, 6, 7], [0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [3, 0, 4, 7]]
TET_FACES = [[0, 2, 1], [0, 1, 3], [1, 2, 3], [0, 3, 2]]


def say(*a):
    print(*a, flush=True)


def boundary_faces(conn):
    """Faces used by one element only, each in its element's local vertex
    order."""
    local = TET_FACES if conn.shape[1] == 4 else HEX_FACES
    seen = {}
    for e, el in enumerate(conn):
        for lf in local:
            f = tuple(int(el[j]) for j in lf)
            key = tuple(sorted(f))
            seen.setdefault(key, []).append(f)
    return np.array([v[0] for v in seen.values() if len(v) == 1], np.int32)


def groups_by_height(X, faces):
    z = X[faces][:, :, 2].mean(axis=1)
    zmax, zmin = X[:, 2].max(), X[:, 2].min()
    g = np.full(len(faces), 2, np.int32)
    g[z > zmax - 1e-9] = 0
    g[z < zmin + 1e-9] = 1
    return g


@pytest.mark.parametrize("element,scheme", [("tet", native.BACKWARD_EULER), ("hex", native.CRANK_NICOLSON)])
@pytest.mark.parametrize("method", [native.MINIMIZE_NEWTON, native.MINIMIZE_NEWTON_LINESEARCH])
def test_pressure_terms_match_oracle(orc, element, scheme, method):
    sc = mesh.beam(nx=8, ny=2, nz=2, element=element, e_range=(2e5, 2e6), seed=3)
    faces = boundary_faces(sc.conn)
    groups = groups_by_height(sc.X, faces)
    p = np.array([800.0, -300.0, 150.0])
    r = native.Rollout.from_scene(sc, method=method, scheme=scheme, precision=native.FP64)
    r.set_pressure(faces, groups)
    r.set_pressure_activation(p)
    omethod = orc.NEWTON if method == native.MINIMIZE_NEWTON else orc.NEWTON_LS
    oscheme = orc.BE if scheme == native.BACKWARD_EULER else orc.CN
    o = orc.System.from_scene(sc, scheme=oscheme, method=omethod)
    o.set_pressure(faces, groups)
    o.set_pressure_activation(p)
    rng = np.random.default_rng(5)
    q = sc.X.reshape(-1) + rng.normal(scale=2e-4, size=3 * sc.n_nodes) * (1 - sc.bcmask)
    dt = sc.dt
    g, go = r.gradient(q, dt), o.gradient(q, dt)
    e, eo = r.energy(q, dt), o.energy(q, dt)
    rp, cols, H = r.hessian(q, dt, precision=native.FP64, project=False)
    orp, ocols, Ho = o.hessian(q, dt, project=False)
    say(element, method, "grad rel", np.abs(g - go).max() / np.abs(go).max(), "energy", e, eo,
        "hess rel", np.abs(H - Ho).max() / np.abs(Ho).max(), "faces", len(faces))
    assert np.abs(g - go).max() <= 1e-12 * np.abs(go).max()
    assert abs(e - eo) <= 1e-12 * abs(eo)
    assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)
    assert np.abs(H - Ho).max() <= 1e-11 * np.abs(Ho).max()
    # without pressure: the gradient changes; the Hessian only under minimize_newton
    r.set_pressure_activation(np.zeros(3))
    assert np.abs(r.gradient(q, dt) - g).max() > 1e-6 * np.abs(g).max()
    H0 = r.hessian(q, dt, precision=native.FP64, project=False)[2]
    assert (np.abs(H0 - H).max() > 0) == (method == native.MINIMIZE_NEWTON)
    r.close()


def reference_cube():
    X = np.array([[0, 0, 0], [0, 1, 0], [1, 1, 0], [1, 0, 0], [0, 0, 1], [0, 1, 1], [1, 1, 1], [1, 0, 1]], float)
    T = np.array([[0, 1, 3, 4], [1, 4, 5, 6], [1, 2, 3, 6], [3, 4, 6, 7], [1, 3, 4, 6]], np.int32)
    mask = np.zeros(24, np.int32)
    mask[:12] = 1   # as written in _test.cpp:333-338 (vertices 0-3)
    return X, T, mask


@pytest.mark.parametrize("precision", [native.FP64, native.MIXED])
def test_reference_pressure_cube_trajectory(orc, precision):
    """_test.cpp:310-382: pressure ramped to 1000 Pa on faces (1, 5, 6) and
    (1, 6, 2), minimize_newton, dt = 0.01: every step lowers the energy at
    its actuation, and the states follow the oracle's (LDL^T Newton)."""
    X, T, mask = reference_cube()
    nel = len(T)
    args = (X, T, np.full(nel, 250000.0), np.full(nel, 0.4), np.full(nel, 1000.0), mask, np.zeros(24))
    r = native.Rollout(*args, grav=(0.0, 0.0, -9.81), method=native.MINIMIZE_NEWTON, precision=precision,
                       predict=0, coarse_nodes=0, g_tol=1e-10, pcg_rtol=1e-10)
    o = orc.System(*args, grav=(0.0, 0.0, -9.81), method=orc.NEWTON)
    faces = np.array([[1, 5, 6], [1, 6, 2]], np.int32)
    r.set_pressure(faces, [0, 0])
    o.set_pressure(faces, [0, 0])
    dt = 0.01
    q = qo = X.reshape(-1).copy()
    for i in range(1, 10):
        p = [min(1000.0 * 2 * i / 10, 1000.0)]
        r.set_pressure_activation(p)
        o.set_pressure_activation(p)
        e0 = r.energy(q, dt)
        q, st = r.step(q, dt)
        qo, _ = o.step(qo, dt)
        e1 = r.energy(q, dt)
        assert e1 <= e0
        err = np.abs(q - qo).max()
        say("cube step", i, "pressure", p[0], "err", err, "newton", st.newton_iters)
        assert err <= (1e-9 if precision == native.FP64 else 1e-6)
    assert np.abs(q - X.reshape(-1)).max() > 1e-4   # the pressure moved it
    r.close()


def load_py_sors():
    sys.path.insert(0, PKG)
    try:
        return importlib.import_module("py_sors")
    finally:
        sys.path.remove(PKG)


def test_sopra_style_pressure_driver(orc):
    """python/d04_sopra.py's structure on a small tet beam: faces from
    IO.extract_tet_surfaces (a solid beam has one surface, its skin: the top
    faces are group 0, the rest group 1, where the demo's chambers would be
    groups of their own); pressure ramped per group through the "pressure"
    actuation, minimize_newton, against the oracle's LDL^T Newton."""
    ps = load_py_sors()
    sc = mesh.beam(nx=10, ny=2, nz=2, element="tet", steps=12)
    surf, _ = ps.IO.extract_tet_surfaces(sc.X, sc.conn)
    faces = np.concatenate([np.array(g, np.int32) for g in surf])
    groups = groups_by_height(sc.X, faces)
    groups[groups == 2] = 1
    settings = ps.SimulationSettings()
    settings.meshType = "Tetrahedron"
    settings.timeSteppingScheme = "backward_euler"
    settings.solverMethod = "minimize_newton"
    settings.dt = sc.dt
    settings.pcgTolerance = 1e-8
    energies = [set(["neohookean"]) for _ in range(sc.n_elems)]
    params = [ps.Params(["density", "youngsModulus", "poissonsRatio"], [1100.0, 1e6, 0.45])
              for _ in range(sc.n_elems)]
    cparams = ps.Params(["neumannBCmask", "neumannBCvalue"], [sc.bcmask.astype(float), sc.bcvalue])
    fparams = ps.Params(["surfaceVertexIdx", "surfaceGroups"], [faces.astype(float), groups.astype(float)])
    energy = ps.Energy(sc.X, sc.conn, settings, sc.grav, 0.0, energies, params, ["neumannBC"], cparams,
                       ["pressure"], fparams)
    solver = ps.Solver(settings)
    o = orc.System.from_scene(sc, method=orc.NEWTON)
    o.set_initial_velocity(np.zeros(3 * sc.n_nodes))
    o.set_pressure(faces, groups)
    act = ps.Params(["pressure"], [np.zeros([2, 1])])
    q = qo = sc.X.reshape(-1).copy()
    for t in range(sc.steps):
        p = np.array([[4000.0 * min(1.0, (t + 1) / 6)], [0.0]])
        act.set_param("pressure", p)
        o.set_pressure_activation(p.reshape(-1))
        q = solver.step(q, energy, act, 0.0)
        qo, _ = o.step(qo, sc.dt)
    err = np.abs(q - qo).max() / sc.bbox_diag
    say("sopra-style driver vs oracle", err, "max displacement", np.abs(q - sc.X.reshape(-1)).max())
    assert err <= 1e-6
    assert np.abs(q - sc.X.reshape(-1)).max() > 1e-5
