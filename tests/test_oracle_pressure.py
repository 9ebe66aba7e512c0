"""PressureForce in the oracle (cpp/src/forces/pressureForce.cpp:14-185),
pinned by a closed form, finite differences and the reference's own test.

- A flat square of two triangles (counter-clockwise seen from +z) under
  pressure p: every vertex of a face takes -p A / 3 along z, the total is
  -p (area) e_z.
- Triangle faces: the Jacobian blocks (j, k) = p/6 hessX[k] equal the
  derivative of -f_j with respect to x_k (central differences).
- Quad faces: the reference keeps p/6 with four vertices per face, so its
  blocks are 4/3 of the derivative of -f (the factor of a consistent
  Jacobian is p/8); the oracle reproduces this as written.
- cpp/src/_test.cpp:310-382 (TestCube::test_energy_min): a five-tet cube with
  pressure ramped to 1000 Pa on two top triangles, minimize_newton: the
  energy at the same actuation decreases over every step.
"""
import numpy as np

from oracle import oracle as O


def square_tets():
    # the reference's cube of five tetrahedra (_test.cpp:290-307)
    X = np.array([[0, 0, 0], [0, 1, 0], [1, 1, 0], [1, 0, 0], [0, 0, 1], [0, 1, 1], [1, 1, 1], [1, 0, 1]], float)
    T = np.array([[0, 1, 3, 4], [1, 4, 5, 6], [1, 2, 3, 6], [3, 4, 6, 7], [1, 3, 4, 6]], np.int32)
    return X, T


def system(X, T, method=O.NEWTON, mask=None, scheme=O.BE):
    n = X.size
    nel = len(T)
    return O.System(X, T, np.full(nel, 250000.0), np.full(nel, 0.4), np.full(nel, 1000.0),
                    bcmask=mask if mask is not None else np.zeros(n, np.int32), bcvalue=np.zeros(n),
                    grav=(0.0, 0.0, -9.81), scheme=scheme, method=method)


def test_flat_square_closed_form():
    X, T = square_tets()
    s = system(X, T)
    # bottom face z = 0: vertices 0 (0,0), 3 (1,0), 2 (1,1), 1 (0,1) counter-clockwise from +z
    faces = np.array([[0, 3, 2], [0, 2, 1]], np.int32)
    s.set_pressure(faces, [0, 0])
    p = 250.0
    s.set_pressure_activation([p])
    f = s.pressure_force(X.reshape(-1)).reshape(-1, 3)
    # each triangle has area 1/2, normal +z: each of its vertices gets -p/6 e_z
    expect = np.zeros((8, 3))
    for face in faces:
        for v in face:
            expect[v, 2] += -p * 0.5 / 3
    assert np.allclose(f, expect, rtol=0, atol=1e-13)
    assert abs(f[:, 2].sum() + p * 1.0) < 1e-12


def fd_jacobian(s, q, faces, h=1e-6):
    """d(-f)/dx by central differences, blocks per face (j, k)."""
    nf, nv = faces.shape
    J = np.zeros((nf, nv, nv, 3, 3))
    for fi, face in enumerate(faces):
        for k, vk in enumerate(face):
            for m in range(3):
                dq = np.zeros_like(q)
                dq[3 * vk + m] = h
                df = -(s.pressure_force(q + dq) - s.pressure_force(q - dq)) / (2 * h)
                for j, vj in enumerate(face):
                    J[fi, j, k, :, m] = df[3 * vj:3 * vj + 3]
    return J


def test_triangle_jacobian_is_the_derivative_of_minus_f():
    X, T = square_tets()
    s = system(X, T)
    faces = np.array([[1, 5, 6]], np.int32)   # one face only, so per-face blocks are the full derivative
    s.set_pressure(faces, [0])
    s.set_pressure_activation([1000.0])
    rng = np.random.default_rng(0)
    q = X.reshape(-1) + rng.normal(scale=0.05, size=X.size)
    J = s.pressure_jacobian(q)
    Jfd = fd_jacobian(s, q, faces)
    assert np.abs(J - Jfd).max() <= 1e-7 * np.abs(Jfd).max()


def test_quad_jacobian_is_four_thirds_of_the_derivative():
    # one hexahedron, its top face as a quad
    X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    H = np.arange(8, dtype=np.int32).reshape(1, 8)
    s = system(X, H)
    faces = np.array([[4, 5, 6, 7]], np.int32)
    s.set_pressure(faces, [0])
    s.set_pressure_activation([700.0])
    rng = np.random.default_rng(1)
    q = X.reshape(-1) + rng.normal(scale=0.05, size=X.size)
    J = s.pressure_jacobian(q)
    Jfd = fd_jacobian(s, q, faces)
    assert np.abs(J - 4.0 / 3.0 * Jfd).max() <= 1e-7 * np.abs(Jfd).max()
    # a flat unit quad: total force -p e_z, a quarter on each vertex
    f = s.pressure_force(X.reshape(-1)).reshape(-1, 3)
    assert np.allclose(f[4:, 2], -700.0 / 4, atol=1e-12) and np.allclose(f[:4], 0.0)


def test_reference_cube_energy_decreases_under_ramped_pressure():
    X, T = square_tets()
    mask = np.zeros(24, np.int32)
    # as written in _test.cpp:333-338: the loop sets the first four vertices'
    # mask entries (3i, 3i+1, 3i+2 for i < 4), not those of {0, 3, 4, 7}
    mask[:12] = 1
    s = system(X, T, method=O.NEWTON, mask=mask)
    s.set_pressure(np.array([[1, 5, 6], [1, 6, 2]], np.int32), [0, 0])
    dt = 0.01
    q = X.reshape(-1).copy()
    for i in range(1, 10):
        s.set_pressure_activation([min(1000.0 * 2 * i / 10, 1000.0)])
        e0 = s.energy(q, dt)
        q, _ = s.step(q, dt)
        e1 = s.energy(q, dt)
        assert e1 <= e0, (i, e0, e1)


def test_pressure_hessian_only_for_minimize_newton():
    X, T = square_tets()
    q = X.reshape(-1) + 0.01 * np.sin(np.arange(24.0))
    vals = {}
    for method in (O.NEWTON, O.NEWTON_LS):
        s = system(X, T, method=method)
        s.set_pressure(np.array([[1, 5, 6], [1, 6, 2]], np.int32), [0, 0])
        s.set_pressure_activation([500.0])
        vals[method] = s.hessian(q, 0.01)
        s.set_pressure_activation([0.0])
        vals[(method, 0)] = s.hessian(q, 0.01)
    # Newton: the Jacobian is in; Newton + line search: it is not
    assert not np.array_equal(vals[O.NEWTON][2], vals[(O.NEWTON, 0)][2])
    assert np.array_equal(vals[O.NEWTON_LS][2], vals[(O.NEWTON_LS, 0)][2])
    # the added part is the sum of the Jacobian blocks (BE: scale 1)
    s = system(X, T, method=O.NEWTON)
    s.set_pressure(np.array([[1, 5, 6], [1, 6, 2]], np.int32), [0, 0])
    s.set_pressure_activation([500.0])
    J = s.pressure_jacobian(q)
    D = np.zeros((24, 24))
    for fi, face in enumerate([[1, 5, 6], [1, 6, 2]]):
        for j, vj in enumerate(face):
            for k, vk in enumerate(face):
                D[3 * vj:3 * vj + 3, 3 * vk:3 * vk + 3] += J[fi, j, k]
    rp, cols, v1 = vals[O.NEWTON]
    _, _, v0 = vals[(O.NEWTON, 0)]
    Hd = np.zeros((24, 24))
    for r in range(24):
        Hd[r, cols[rp[r]:rp[r + 1]]] = v1[rp[r]:rp[r + 1]] - v0[rp[r]:rp[r + 1]]
    assert np.abs(Hd - D).max() <= 1e-12 * np.abs(D).max()
