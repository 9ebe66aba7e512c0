"""Pins the oracle's contact terms (penalty + lagged Coulomb friction).

The reference resolves contact as hard constraints (SQP/OSQP); the penalty
model is this build's formulation.  What pins it:
  - the reference's own test scenario cpp/src/_test.cpp:629-698 (a tet with
    one vertex 0.1 m below the plane z = 0; after one backward Euler step no
    vertex is below the plane by more than 1e-4);
  - the closed-form single-vertex penalty force k d n and the geometry of
    Plane3D / Disk3D (plane3D.h:30-33, disk3D.h:33-45);
  - central finite differences of energy -> gradient and gradient -> Hessian
    for each branch (plane, disk interior, disk rim, friction stick/slip);
    the node block equals the PSD projection of the finite-difference block;
  - the friction smoothing's closed form (f0 of Li et al. 2020): full
    Coulomb force mu * N when sliding, never more than mu * N.
"""
import numpy as np
import pytest

from sors_amd import mesh

TET_X = np.array([[0.0, 0.0, -0.10], [1.0, 0.0, 0.20], [0.0, 1.0, 0.20], [0.0, 0.0, 0.30]])
TET_CONN = np.array([[0, 1, 2, 3]], np.int32)


def one_node(orc, **kw):
    """A single tet whose node 0 is the only contact candidate; mass-free so
    the system terms at dt = 0 are the elastic + contact terms only."""
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) + np.array([0.3, 0.2, 0.5])
    s = orc.System(X, TET_CONN, [0.0], [0.0], [0.0], method=orc.NEWTON)
    s.set_contact(surface=[0], **kw)
    return s, X


def test_reference_plane_contact_no_penetration(orc):
    # _test.cpp:629-698: density 1, E 1e5, nu 0.3, gravity -z, dt 1e-2, plane z = 0
    # (normal +z), all four vertices are surface vertices.  A penalty of
    # 1e7 N/m stands in for the hard constraint.
    s = orc.System(TET_X, TET_CONN, [1e5], [0.3], [1.0], grav=(0, 0, -9.81), method=orc.NEWTON_LS)
    s.set_contact(planes=[((0, 0, 0), (0, 0, 1))], surface=[0, 1, 2, 3], stiffness=1e7)
    q, st = s.step(TET_X.reshape(-1), 1e-2)
    z = q.reshape(-1, 3)[:, 2]
    assert z.min() >= -1e-4, z
    # and without contact the vertex stays below
    s2 = orc.System(TET_X, TET_CONN, [1e5], [0.3], [1.0], grav=(0, 0, -9.81), method=orc.NEWTON_LS)
    q2, _ = s2.step(TET_X.reshape(-1), 1e-2)
    assert q2.reshape(-1, 3)[0, 2] < -0.09


def test_single_vertex_penalty_force_closed_form(orc):
    k = 2e5
    n = np.array([1.0, 2.0, 2.0]) / 3.0
    p = np.array([0.3, 0.2, 0.5])
    s, X = one_node(orc, planes=[(p, n * 7.0)], stiffness=k)  # normal is normalised (Plane3D ctor)
    for d in (-3e-3, -1e-6, 2e-3):
        q = X.copy()
        q[0] = p + d * n
        g = s.gradient(q.reshape(-1), 0.0)  # elastic part is zero: q[1:] = X[1:] but q[0] moved
        ge = np.zeros(12)
        s0, _ = one_node(orc)  # no primitives: elastic gradient alone
        ge = s0.gradient(q.reshape(-1), 0.0)
        fc = g - ge
        expect = k * d * n if d < 0 else np.zeros(3)
        assert np.allclose(fc[:3], expect, rtol=1e-12, atol=1e-12 * k * 1e-3)
        assert np.all(fc[3:] == 0.0)
        # d is recomputed from coordinates of size 0.5: rounding of 1e-16 on d
        assert s.contact_energy(q.reshape(-1), 0.0) == pytest.approx(0.5 * k * min(d, 0.0) ** 2, rel=1e-8, abs=0)


def test_disk_geometry_closed_form(orc):
    # disk normal -z at height 0.5 above node 0's xy; radius 0.1
    k = 1e5
    c = np.array([0.3, 0.2, 0.5])
    s, X = one_node(orc, disks=[(c, (0, 0, -1.0), 0.1)], stiffness=k)
    q = X.copy()
    for (dx, dz, w) in [(0.05, 1e-3, 1.0), (0.0999, 2e-3, 1.0), (0.2, 1e-3, 0.0), (0.05, -1e-3, 0.0)]:
        q[0] = c + np.array([dx, 0.0, dz])       # dz > 0: above the disk plane = penetrating (d = -dz)
        e = s.contact_energy(q.reshape(-1), 0.0)
        assert e == pytest.approx(0.5 * k * w * max(dz, 0.0) ** 2, rel=1e-12, abs=1e-300)
    # rim: smooth cubic falloff over 1 mm
    q[0] = c + np.array([0.1005, 0.0, 1e-3])
    t = 0.5
    assert s.contact_energy(q.reshape(-1), 0.0) == pytest.approx(0.5 * k * (1 - 3 * t * t + 2 * t**3) * 1e-6,
                                                                 rel=1e-9)


def fd_grad(f, x, h):
    g = np.zeros_like(x)
    for i in range(x.size):
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        g[i] = (f(xp) - f(xm)) / (2 * h)
    return g


@pytest.mark.parametrize("case", ["plane", "disk_inside", "disk_rim"])
def test_penalty_fd(orc, case):
    k = 3e5
    c = np.array([0.3, 0.2, 0.5])
    if case == "plane":
        kw = dict(planes=[(c, (0.2, -0.1, 1.0))])
        off = np.array([0.01, -0.02, -2e-3])
    else:
        kw = dict(disks=[(c, (0.0, 0.1, -1.0), 0.05)])
        off = np.array([0.02, 0.01, 2e-3]) if case == "disk_inside" else np.array([0.0504, 0.0, 2e-3])
    s, X = one_node(orc, stiffness=k, **kw)
    s0, _ = one_node(orc)
    x0 = c + off
    q = X.reshape(-1).copy()

    def energy(x):
        qq = q.copy()
        qq[:3] = x
        return s.contact_energy(qq, 0.0)

    def grad(x):
        qq = q.copy()
        qq[:3] = x
        return (s.gradient(qq, 0.0) - s0.gradient(qq, 0.0))[:3]

    g = grad(x0)
    assert np.allclose(g, fd_grad(energy, x0, 1e-7), rtol=1e-6, atol=1e-7 * np.abs(g).max())
    qq = q.copy()
    qq[:3] = x0
    H = s.hessian_dense(qq, 0.0)[:3, :3] - s0.hessian_dense(qq, 0.0)[:3, :3]
    assert np.all(np.linalg.eigvalsh(0.5 * (H + H.T)) >= -1e-9 * k)
    Hfd = np.stack([fd_grad(lambda x: grad(x)[i], x0, 1e-8) for i in range(3)])
    Hfd = 0.5 * (Hfd + Hfd.T)
    lam, V = np.linalg.eigh(Hfd)
    Hpsd = (V * np.maximum(lam, 0.0)) @ V.T
    if case == "disk_rim":
        say_rim = np.linalg.eigvalsh(Hfd)
        assert say_rim.min() < 0  # this point exercises the projection
    assert np.allclose(H, Hpsd, rtol=1e-6, atol=1e-6 * np.abs(Hfd).max())


@pytest.mark.parametrize("slide", [2e-9, 5e-7, 5e-5])  # stick (|uT| < eps = 1e-6), transition, slip
def test_friction_fd_and_coulomb_bound(orc, slide):
    """Node 0 rests 1 um inside the plane z = 0.5 + 1e-6 (normal +z); the lag
    gives N = k * 1e-6.  Move it tangentially by `slide` over one step."""
    k, mu, eps_v, dt = 1e5, 0.5, 1e-3, 1e-3
    depth = 1e-6
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) + np.array([0.3, 0.2, 0.5])
    s = orc.System(X, TET_CONN, [0.0], [0.0], [0.0], method=orc.NEWTON)
    s0 = orc.System(X, TET_CONN, [0.0], [0.0], [0.0], method=orc.NEWTON)
    s.set_contact(planes=[((0, 0, 0.5 + depth), (0, 0, 1))], surface=[0], stiffness=k, friction=mu, eps_v=eps_v)
    s.contact_lag(dt)  # from qprev = X
    N = k * depth
    eps = eps_v * dt
    q = X.reshape(-1).copy()
    x0 = X[0] + np.array([slide * 0.6, slide * 0.8, 0.0])

    def energy(x):
        qq = q.copy()
        qq[:3] = x
        return s.contact_energy(qq, dt)

    def grad(x):
        qq = q.copy()
        qq[:3] = x
        return (s.gradient(qq, dt) - s0.gradient(qq, dt))[:3]

    g = grad(x0)
    assert g[2] == pytest.approx(-k * depth, rel=1e-8)  # the penalty; friction is tangential (n = +z)
    ft = np.array([g[0], g[1], 0.0])
    assert np.linalg.norm(ft) <= mu * N * (1 + 1e-9)  # N from a depth recovered to 1e-10 relative
    y = slide
    f1 = 1.0 if y >= eps else -y * y / eps**2 + 2 * y / eps  # f0' (Li et al. 2020)
    # y is recovered from coordinates of size 0.3: rounding 1e-16 / 2e-9
    assert np.linalg.norm(ft) == pytest.approx(mu * N * f1, rel=1e-6)
    assert np.allclose(ft / np.linalg.norm(ft), [0.6, 0.8, 0.0], atol=1e-6)
    h = min(1e-10, 1e-2 * slide)  # inside one branch, small against eps
    assert np.allclose(g, fd_grad(energy, x0, h), rtol=1e-6, atol=1e-6 * np.abs(g).max())
    qq = q.copy()
    qq[:3] = x0
    H = s.hessian_dense(qq, dt)[:3, :3] - s0.hessian_dense(qq, dt)[:3, :3]
    Hfd = np.stack([fd_grad(lambda x: grad(x)[i], x0, h) for i in range(3)])
    assert np.allclose(H, Hfd, rtol=1e-5, atol=1e-5 * np.abs(H).max())
    assert np.all(np.linalg.eigvalsh(0.5 * (H + H.T)) >= -1e-9 * np.abs(H).max())


def test_cube_rests_on_ground_and_friction_holds_slope(orc):
    """A small cube on a 20-degree slope: with mu = 0.5 > tan(20 deg) it
    stays (creeps by less than the stick tolerance); frictionless it slides."""
    X, hexes = mesh.box_grid(2, 2, 2, 0.02, 0.02, 0.02)
    conn = mesh.kuhn_split(hexes).astype(np.int32)
    th = np.deg2rad(20.0)
    nrm = np.array([0.0, -np.sin(th), np.cos(th)])
    R = np.array([[1, 0, 0], [0, np.cos(th), -np.sin(th)], [0, np.sin(th), np.cos(th)]])
    Xr = X @ R.T
    Xr -= nrm * (Xr @ nrm).min()  # bottom face touching the plane through the origin
    bottom = np.nonzero(np.abs(Xr @ nrm) < 1e-9)[0]
    E = np.full(conn.shape[0], 1e5)
    out = {}
    for mu in (0.0, 0.5):
        s = orc.System(Xr, conn, E, np.full(len(E), 0.3), np.full(len(E), 1000.0), grav=(0, 0, -9.81),
                       method=orc.NEWTON_LS)
        s.set_contact(planes=[((0, 0, 0), nrm)], stiffness=1e5, friction=mu, eps_v=1e-4)
        q = Xr.reshape(-1).copy()
        for _ in range(30):
            q, _ = s.step(q, 1e-3)
        d = q.reshape(-1, 3) - Xr
        out[mu] = np.linalg.norm(d[bottom] - (d[bottom] @ nrm)[:, None] * nrm, axis=1).max()
        # no node sinks more than the penalty's static depth (weight / (k * contact nodes))
        assert (q.reshape(-1, 3) @ nrm).min() > -1e-4
    assert out[0.0] > 1e-3              # frictionless: slides down the slope
    assert out[0.5] < 0.05 * out[0.0]   # Coulomb: held
