"""The oracle's minimize_SQP (solver.cpp:109-197) and its QP solver.

The reference hands each SQP iteration's QP to OSQP, which is absent here
(empty submodule external/osqp); the oracle restates OSQP's ADMM with its
solution polishing.  What pins it:
  - brute force: every active subset of tiny QPs solved as a KKT system; the
    one feasible point with multipliers of the right sign is the solution;
  - closed forms: no rows (a Newton step), one row (projection in the
    P-metric);
  - the reference's own contact test (_test.cpp:629-698): a tetrahedron with
    a vertex 0.1 m below the plane z = 0 ends one SQP step with no
    penetration beyond TOL = 1e-4 (_test.h:210);
  - the active-set rules of PlaneContact / DiskContact on geometry whose
    answer is fixed by hand (planeContact.cpp:87-124, diskContact.cpp:91-202);
  - the converged SQP state satisfies the KKT conditions of the step
    problem; with only NeumannBC rows SQP and minimize_newton (its masked
    Dirichlet handling) reach the same minimiser.
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as O


def brute_force_qp(P, q, A, l, u):
    """Enumerate the active subsets of the inequality rows."""
    m, n = A.shape
    eq = [i for i in range(m) if l[i] == u[i]]
    ineq = [i for i in range(m) if l[i] != u[i]]
    found = []
    for k in range(len(ineq) + 1):
        for S in itertools.combinations(ineq, k):
            act = eq + list(S)
            Aa = A[act]
            K = np.block([[P, Aa.T], [Aa, np.zeros((len(act), len(act)))]])
            rhs = np.concatenate([-q, l[act]])
            try:
                sol = np.linalg.solve(K, rhs)
            except np.linalg.LinAlgError:
                continue
            x, y = sol[:n], sol[n:]
            if np.all(A @ x >= l - 1e-10) and np.all(y[len(eq):] <= 1e-10):
                found.append(x)
    assert found, "no KKT point"
    for x in found[1:]:
        assert np.allclose(x, found[0], atol=1e-9)
    return found[0]


def dense_csr(n):
    return np.arange(0, n * n + 1, n), np.tile(np.arange(n), n)


def random_qp(seed, nv=3, mmax=5):
    rng = np.random.default_rng(seed)
    n = 3 * nv
    B = rng.standard_normal((n, n))
    P = B @ B.T + 0.5 * np.eye(n)
    q = rng.standard_normal(n)
    m = int(rng.integers(1, mmax + 1))
    vtx, cnt = [], [0] * nv
    while len(vtx) < m:  # at most two rows per vertex: the rows stay independent
        v = int(rng.integers(nv))
        if cnt[v] < 2:
            cnt[v] += 1
            vtx.append(v)
    a = rng.standard_normal((m, 3))
    A = np.zeros((m, n))
    for i in range(m):
        A[i, 3 * vtx[i]:3 * vtx[i] + 3] = a[i]
    x0 = rng.standard_normal(n)  # a feasible point: every row holds at x0
    slack = np.where(rng.random(m) < 0.3, 0.0, rng.random(m))
    l = A @ x0 - slack
    u = np.full(m, np.inf)
    iseq = rng.random(m) < 0.25
    l[iseq] = (A @ x0)[iseq]
    u[iseq] = l[iseq]
    return P, q, A, vtx, a, l, u


@pytest.mark.parametrize("seed", range(40))
def test_qp_matches_brute_force_enumeration(seed):
    P, q, A, vtx, a, l, u = random_qp(seed)
    rp, cols = dense_csr(P.shape[0])
    x, y, it = O.qp_solve(rp, cols, P.reshape(-1), q, vtx, a, l, u)
    xb = brute_force_qp(P, q, A, l, u)
    assert np.abs(x - xb).max() <= 1e-9 * max(1.0, np.abs(xb).max())
    # OSQP's multiplier convention: P x + q + A^T y = 0, y <= 0 on an active lower bound
    assert np.abs(P @ x + q + A.T @ y).max() <= 1e-8 * max(1.0, np.abs(q).max())
    ineq = u == np.inf
    assert np.all(y[ineq] <= 1e-9)
    assert np.all(np.abs(y[ineq] * (A @ x - l)[ineq]) <= 1e-8)   # complementarity


def test_qp_without_rows_is_the_newton_step():
    P, q, *_ = random_qp(7)
    rp, cols = dense_csr(P.shape[0])
    x, y, _ = O.qp_solve(rp, cols, P.reshape(-1), q, [], np.zeros((0, 3)), [], [])
    assert np.allclose(x, np.linalg.solve(P, -q), rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("offset", [-0.5, 0.5])
def test_qp_one_row_is_the_metric_projection(offset):
    """One row a.x_v >= l: inactive when the Newton step satisfies it, else
    x = x_u + P^-1 a' (l - a'.x_u) / (a'^T P^-1 a')."""
    rng = np.random.default_rng(3)
    n = 6
    B = rng.standard_normal((n, n))
    P = B @ B.T + np.eye(n)
    q = rng.standard_normal(n)
    xu = np.linalg.solve(P, -q)
    a = np.array([0.3, -0.2, 0.9])
    ar = np.zeros(n)
    ar[3:6] = a
    l = ar @ xu + offset
    rp, cols = dense_csr(n)
    x, y, _ = O.qp_solve(rp, cols, P.reshape(-1), q, [1], a[None], [l], [np.inf])
    if offset < 0:
        expect = xu
    else:
        Pa = np.linalg.solve(P, ar)
        expect = xu + Pa * (l - ar @ xu) / (ar @ Pa)
    assert np.allclose(x, expect, rtol=1e-10, atol=1e-12)


def test_reference_plane_contact_no_penetration():
    """cpp/src/_test.cpp:629-698: one tetrahedron, vertex 0 at z = -0.1 below
    the plane z = 0 (normal +z), all four vertices on the surface, gravity
    -9.81, E = 1e5, nu = 0.3, density 1, dt = 1e-2, minimize_SQP; after one
    step n.(x - p) >= -TOL with TOL = 1e-4 (cpp/headers/_test.h:210)."""
    X = np.array([[0, 0, -0.10], [1, 0, 0.20], [0, 1, 0.20], [0, 0, 0.30]], float)
    s = O.System(X, [[0, 1, 2, 3]], [1e5], [0.3], [1.0], grav=(0, 0, -9.81), method=O.SQP)
    s.set_contact(planes=[((0, 0, 0), (0, 0, 1))], surface=[0, 1, 2, 3])
    q, st = s.step(X.reshape(-1), 1e-2)
    assert q.reshape(-1, 3)[:, 2].min() >= -1e-4
    assert st.newton_iters >= 2
    # the penetrating vertex is pushed exactly onto the plane (its row active)
    assert abs(q[2]) < 1e-12


def test_plane_active_set_rule():
    """planeContact.cpp:87-124: the plane of smallest signed distance below
    0.01 m, else -1 (planes z = 0 and x = 0.5 facing -x)."""
    X = np.array([[0.1, 0.0, 0.005],    # 5 mm above z = 0, 0.4 m from x = 0.5: plane 0
                  [0.495, 0.0, 0.2],    # 5 mm before x = 0.5: plane 1
                  [0.499, 0.0, 0.004],  # both within 0.01: plane 1 at 1 mm (smaller)
                  [0.2, 0.0, 0.02],     # 2 cm above: none
                  [0.2, 0.0, -0.3]])    # far below z = 0: plane 0 (signed distance -0.3)
    s = O.System(X, [[0, 1, 2, 3]], [1e5], [0.3], [1.0], method=O.SQP)
    s.set_contact(planes=[((0, 0, 0), (0, 0, 1)), ((0.5, 0, 0), (-1, 0, 0))], surface=[0, 1, 2, 3, 4])
    pm, dm, fr = s.sqp_active(X.reshape(-1))
    assert list(pm) == [0, 1, 1, -1, 0]
    assert list(dm) == [-1] * 5 and fr == 0


def grid_points(z=0.0):
    g = np.arange(-2, 3) * 0.004
    return np.array([[x, y, z] for x in g for y in g])


def test_disk_active_set_capture_freeze_release():
    """diskContact.cpp:91-202 on a 5 x 5 grid (4 mm pitch) at z = 0 under a
    disk of radius 5 mm facing -z: capture within radius + 1 mm (9 points),
    frozen while the disk is less than 3 x 0.01 m above, released beyond."""
    X = grid_points()
    conn = [[0, 1, 5, 6]]
    s = O.System(X, conn, [1e5], [0.3], [1.0], method=O.SQP)
    surf = list(range(len(X)))
    s.set_contact(disks=[((0, 0, 0.008), (0, 0, -1), 0.005)], surface=surf)
    x = X.reshape(-1)
    pm, dm, fr = s.sqp_active(x)
    inside = np.hypot(X[:, 0], X[:, 1]) <= 0.006
    assert inside.sum() == 9
    assert fr == 1 and np.array_equal(dm == 0, inside)
    # disk lifted to 25 mm (< 30 mm clear distance): the footprint stays frozen
    s.set_disk(0, (0, 0, 0.025), (0, 0, -1), 0.005)
    _, dm2, fr2 = s.sqp_active(x)
    assert fr2 == 1 and np.array_equal(dm2, dm)
    # beyond 30 mm above every frozen vertex: released, nothing captured in that call
    s.set_disk(0, (0, 0, 0.031), (0, 0, -1), 0.005)
    _, dm3, fr3 = s.sqp_active(x)
    assert fr3 == 0 and np.all(dm3 == -1)
    # a disk 2 cm away (no vertex within 1 cm of its plane): nothing captured
    s.set_disk(0, (0, 0, 0.02), (0, 0, -1), 0.005)
    _, dm4, fr4 = s.sqp_active(x)
    assert fr4 == 0 and np.all(dm4 == -1)
    # a vertex 6 cm beyond the disk plane is outside the capture window (-5 thresholds)
    X2 = X.copy()
    X2[12, 2] = 0.07
    s.set_disk(0, (0, 0, 0.008), (0, 0, -1), 0.005)
    _, dm5, fr5 = s.sqp_active(X2.reshape(-1))
    assert fr5 == 1 and dm5[12] == -1 and (dm5 == 0).sum() == 8


def cube_on_plane(n=3, h=0.02, lift=-1e-3, v0z=-0.2):
    """A Kuhn-tet cube of side h resting on z = 0 (lifted by `lift`), moving down."""
    from sors_amd import mesh

    X, hexes = mesh.box_grid(n, n, n, h, h, h)
    conn = mesh.kuhn_split(hexes)
    surf = mesh.box_surface_nodes(X)
    X[:, 2] += lift
    return X, conn, surf, v0z


def test_sqp_converged_state_is_a_kkt_point():
    """A tet cube pressed into the ground plane: after each SQP step the
    state satisfies the step problem's KKT conditions with the active set of
    its start (stationarity with nonnegative contact forces, feasibility,
    complementarity)."""
    X, conn, surf, v0z = cube_on_plane()
    N = len(X)
    s = O.System(X, conn, np.full(len(conn), 2e5), np.full(len(conn), 0.4), np.full(len(conn), 1000.0),
                 grav=(0, 0, -9.81), method=O.SQP)
    v0 = np.zeros((N, 3))
    v0[:, 2] = v0z
    s.set_initial_velocity(v0.reshape(-1))
    s.set_contact(planes=[((0, 0, 0), (0, 0, 1))], surface=surf)
    q = X.reshape(-1).copy()
    dt = 5e-3
    for step in range(4):
        pm, _, _ = s.sqp_active(q)   # the active set the step will use (plane rule has no memory)
        q, st = s.minimize(q, dt)    # the step's minimiser, before the state update
        g = s.gradient(q, dt)
        z = q.reshape(-1, 3)[:, 2]
        act = np.array(surf)[pm >= 0]
        assert z[act].min() >= -1e-12                       # feasibility of the active rows
        lam = g.reshape(-1, 3)[act, 2]                      # g = A^T lambda on the rows (normal +z)
        r = g.reshape(-1, 3).copy()
        r[act, 2] -= lam
        tol = 1e-6 * np.abs(g).max() + 1e-10                # N (the cube may bounce off: g -> 0)
        assert np.abs(r).max() <= tol                       # stationarity
        assert lam.min() >= -tol                            # forces push outward
        assert np.abs(lam * z[act]).max() <= 1e-3 * tol     # complementarity (z in m)
        s.update_state(q, dt)
    assert st.newton_iters >= 1


def test_sqp_with_only_dirichlet_rows_matches_newton():
    """NeumannBC rows alone (no contact): SQP's equality rows and
    minimize_newton's masked Dirichlet handling give the same minimiser."""
    from sors_amd import mesh

    sc = mesh.beam(nx=6, ny=2, nz=2, steps=3)
    a = O.System.from_scene(sc, method=O.SQP)
    b = O.System.from_scene(sc, method=O.NEWTON)
    qa = qb = sc.X.reshape(-1).copy()
    for _ in range(sc.steps):
        qa, _ = a.step(qa, sc.dt)
        qb, _ = b.step(qb, sc.dt)
    assert np.abs(qa - qb).max() <= 1e-9 * sc.bbox_diag
