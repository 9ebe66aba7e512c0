"""Pins the oracle's CFL time step (Solver::calculate_CFL_timestep,
solver.cpp:300-331; Element::compute_minimal_edge_length, element.h:131-142)
against closed forms: the shortest vertex-pair distance of known elements,
including a sheared hex whose face diagonal is shorter than every edge (the
reference takes every vertex pair, not the edges only), the largest |v|
component, the cap at dt and the Courant factor."""
import numpy as np
import pytest

# a cube of side 0.5 (hex8 corner order of the reference, hexahedron.cpp:171-179)
CUBE = 0.5 * np.array([[1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 0], [1, 0, 1], [1, 1, 1], [0, 1, 1], [0, 0, 1]], float)
# right-corner tet with legs 1, 2, 3: pair distances 1, 2, 3, sqrt 5, sqrt 10, sqrt 13
TET = np.array([[0, 0, 0], [1, 0, 0], [0, 2, 0], [0, 0, 3]], float)


def sheared_hex():
    # bottom face A(0,0) B(1,0) C(1.9,0.2) D(0.9,0.2), extruded by 1 in z:
    # edges 1 and sqrt(0.85); face diagonal BD = sqrt(0.01 + 0.04)
    f = np.array([[0, 0], [1, 0], [1.9, 0.2], [0.9, 0.2]])
    X = np.zeros((8, 3))
    X[:4, :2] = f
    X[4:, :2] = f
    X[4:, 2] = 1.0
    return X, np.sqrt(0.05)


def velocity(N, k, val):
    v = np.zeros(3 * N)
    v[k] = val
    return v


def test_cube_hex(orc):
    conn = np.arange(8, dtype=np.int32)[None]
    v = velocity(8, 13, -3.0)  # max |v| component 3
    assert orc.cfl_timestep(CUBE, conn, v, dt=1.0) == pytest.approx(0.5 / 3.0, rel=1e-15)
    assert orc.cfl_timestep(CUBE, conn, v, dt=1.0, cfl=0.5) == pytest.approx(0.25 / 3.0, rel=1e-15)
    assert orc.cfl_timestep(CUBE, conn, v, dt=0.1) == 0.1  # capped at dt


def test_tet_shortest_pair(orc):
    conn = np.arange(4, dtype=np.int32)[None]
    v = velocity(4, 2, 0.25)
    assert orc.cfl_timestep(TET, conn, v, dt=10.0) == pytest.approx(1.0 / 0.25, rel=1e-15)
    # the positions passed in count, not the rest shape: node 1 moved to
    # (3, 0, 0) gives pairs 3, 2, 3, sqrt 13, sqrt 18, sqrt 13
    q = TET.copy()
    q[1] = [3.0, 0, 0]
    assert orc.cfl_timestep(q, conn, v, dt=10.0) == pytest.approx(2.0 / 0.25, rel=1e-15)


def test_sheared_hex_uses_every_vertex_pair(orc):
    X, dmin = sheared_hex()
    conn = np.arange(8, dtype=np.int32)[None]
    v = velocity(8, 0, 1.0)
    assert orc.cfl_timestep(X, conn, v, dt=1.0) == pytest.approx(dmin, rel=1e-15)
    assert dmin < np.sqrt(0.85) < 1.0  # shorter than every edge


def test_no_motion_gives_dt(orc):
    conn = np.arange(8, dtype=np.int32)[None]
    assert orc.cfl_timestep(CUBE, conn, np.zeros(24), dt=0.01) == 0.01


def test_system_uses_its_velocity(orc):
    conn = np.arange(4, dtype=np.int32)[None]
    s = orc.System(TET, conn, [1e5], [0.3], [1000.0])
    s.set_initial_velocity(velocity(4, 5, -0.5))
    assert s.cfl_timestep(TET.reshape(-1), dt=10.0) == pytest.approx(1.0 / 0.5, rel=1e-15)
