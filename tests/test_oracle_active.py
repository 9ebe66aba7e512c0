"""Pins the oracle's active fibre stress (tet4):
  Psi_act = 1/2 sigma a (|F f|^2 - 1) per unit rest volume.
Closed forms: zero energy and force-free only for a = 0; at rest the active
first Piola stress is sigma a f f^T (the config's "sigma_max a (f x f)"), so
the nodal forces are V0 sigma a f (f . g_k); the tangent is
V0 sigma a (f . g_a)(f . g_b) I.  Finite differences pin the rest."""
import numpy as np
import pytest

X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) * 0.01 + 0.1
CONN = np.array([[0, 1, 2, 3]], np.int32)


def sys1(orc, E=0.0, a=0.7, sigma=2e5, f=(0.3, -0.2, 1.0)):
    s = orc.System(X, CONN, [E], [0.4], [0.0], method=orc.NEWTON)
    s.set_active([0], np.array([f], float), sigma)
    s.set_activation([a])
    return s


def test_rest_state_closed_form(orc):
    sigma, a = 2e5, 0.7
    f = np.array([0.3, -0.2, 1.0])
    f /= np.linalg.norm(f)
    s = sys1(orc, a=a, sigma=sigma, f=f)
    q = X.reshape(-1)
    # g_k: rows of Dm^-1, Dm = [X0-X3, X1-X3, X2-X3]
    Dm = (X[:3] - X[3]).T
    G = np.linalg.inv(Dm)
    g = np.vstack([G, -G.sum(0)])
    V0 = abs(np.linalg.det(Dm)) / 6
    expect = np.concatenate([V0 * sigma * a * f * (f @ g[k]) for k in range(4)])
    assert np.allclose(s.gradient(q, 0.0), expect, rtol=1e-12, atol=1e-12 * np.abs(expect).max())
    assert abs(s.energy(q, 0.0)) < 1e-12 * V0 * sigma
    H = s.hessian_dense(q, 0.0)
    Hexp = np.zeros((12, 12))
    for i in range(4):
        for j in range(4):
            Hexp[3 * i:3 * i + 3, 3 * j:3 * j + 3] = V0 * sigma * a * (f @ g[i]) * (f @ g[j]) * np.eye(3)
    assert np.allclose(H, Hexp, rtol=1e-12, atol=1e-12 * np.abs(Hexp).max())
    # passive group (a = 0): no force
    s.set_activation([0.0])
    assert np.abs(s.gradient(q, 0.0)).max() == 0.0


def test_active_plus_neohookean_fd(orc):
    s = sys1(orc, E=5e5, a=0.4)
    rng = np.random.default_rng(3)
    q0 = X.reshape(-1) + rng.normal(scale=5e-4, size=12)
    g = s.gradient(q0, 0.0)
    h = 1e-8
    gfd = np.array([(s.energy(q0 + h * e, 0.0) - s.energy(q0 - h * e, 0.0)) / (2 * h) for e in np.eye(12)])
    assert np.allclose(g, gfd, rtol=1e-6, atol=1e-7 * np.abs(g).max())
    H = s.hessian_dense(q0, 0.0)
    Hfd = np.stack([(s.gradient(q0 + h * e, 0.0) - s.gradient(q0 - h * e, 0.0)) / (2 * h) for e in np.eye(12)])
    assert np.allclose(H, Hfd, rtol=1e-5, atol=1e-6 * np.abs(H).max())


def test_youngs_update_matches_fresh_system(orc):
    s = sys1(orc, E=5e5, a=0.4)
    s.set_youngs([7e5])
    t = sys1(orc, E=7e5, a=0.4)
    rng = np.random.default_rng(4)
    q = X.reshape(-1) + rng.normal(scale=5e-4, size=12)
    assert np.array_equal(s.gradient(q, 0.0), t.gradient(q, 0.0))


def test_activation_contracts_fibre(orc):
    """A free tet with a strongly active fibre along z shortens along z over a
    few implicit steps (the reference's muscle test: activation shortens the
    z-extent, cpp/src/_test.cpp:563-625)."""
    s = orc.System(X, CONN, [1e5], [0.4], [1000.0], method=orc.NEWTON_LS)
    s.set_active([0], np.array([[0.0, 0.0, 1.0]]), 2e5)
    s.set_activation([1.0])
    q = X.reshape(-1).copy()
    for _ in range(20):
        q, _ = s.step(q, 1e-3)
    z = q.reshape(-1, 3)[:, 2]
    assert np.ptp(z) < 0.95 * np.ptp(X[:, 2])
