"""Pins the CPU oracle's element level against the reference's known-answer
tests, torch FP64 autograd golden vectors and closed forms."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "element_nh.npz")

UNIT_TET = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)


def test_unit_tet_deformation_hessian_known_answer(orc):
    # cpp/src/_test.cpp:217-235 (TestTetrahedron::test_deformationHessian)
    expected = np.array([
        [-1, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0],
        [-1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0],
        [-1, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0],
        [0, -1, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0],
        [0, -1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0],
        [0, -1, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0],
        [0, 0, -1, 0, 0, 1, 0, 0, 0, 0, 0, 0],
        [0, 0, -1, 0, 0, 0, 0, 0, 1, 0, 0, 0],
        [0, 0, -1, 0, 0, 0, 0, 0, 0, 0, 0, 1]], float)
    _, V0, B = orc.tet_precompute(UNIT_TET)
    # The reference's test builds the tet with undeformedVertices_ listed
    # as (0,0,0),(1,0,0),(0,1,0),(0,0,1); its literal is for that order.
    assert np.allclose(B, expected)
    assert V0 == pytest.approx(1 / 6)


def test_energy_and_gradient_zero_at_rest(orc):
    # cpp/src/_test.cpp:238-286
    e, g, _ = orc.elem_eval(UNIT_TET, UNIT_TET, 1.0, 0.3)
    assert abs(e) < 1e-9
    assert np.allclose(g, 0, atol=1e-12)
    rng = np.random.default_rng(1)
    X = rng.normal(size=(8, 3)) * 0.01 + np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                                                      [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
    e, g, _ = orc.elem_eval(X, X, 1e6, 0.45)
    assert abs(e) < 1e-9 and np.allclose(g, 0, atol=1e-9)


def test_tet_precompute_inverse_and_volume(orc):
    rng = np.random.default_rng(2)
    for _ in range(10):
        X = rng.normal(size=(4, 3))
        Dminv, V0, _ = orc.tet_precompute(X)
        Dm = np.stack([X[0] - X[3], X[1] - X[3], X[2] - X[3]], axis=1)
        assert np.allclose(Dm @ Dminv, np.eye(3), atol=1e-12)
        assert V0 == pytest.approx(abs(np.linalg.det(Dm)) / 6, rel=1e-13)


def test_hex_precompute_closed_forms(orc):
    # axis-aligned box: shape gradients are analytic, V0 is the box volume,
    # and an affine map x = A X + b gives F = A at every Gauss point.
    h = np.array([0.004, 0.003, 0.005])
    corners = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    X = corners * h
    G, B, V0 = orc.hex_precompute(X)
    assert V0 == pytest.approx(np.prod(h), rel=1e-13)
    g = 1 / np.sqrt(3)
    # the 8 Gauss points in the reference's order, mapped to [0,1]^3 of the box
    xi_ref = np.array([[1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, -1], [1, -1, 1], [1, 1, 1], [-1, 1, 1],
                       [-1, -1, 1]], float)
    for s in range(8):
        # physical point: the reference's reference-cube -> voxel map is affine
        # (node i of the reference cube maps to voxel corner i)
        A = np.linalg.lstsq(np.c_[xi_ref, np.ones(8)], corners, rcond=None)[0]
        u = np.r_[g * xi_ref[s], 1] @ A  # local coordinate in [0,1]^3
        for a in range(8):
            c = corners[a]
            w = [(u[d] if c[d] else 1 - u[d]) for d in range(3)]
            dw = [(1.0 if c[d] else -1.0) for d in range(3)]
            grad = np.array([dw[0] * w[1] * w[2], w[0] * dw[1] * w[2], w[0] * w[1] * dw[2]]) / h
            assert np.allclose(G[s, a], grad, rtol=1e-12, atol=1e-9)
    rng = np.random.default_rng(3)
    Xd = X + rng.normal(scale=3e-4, size=X.shape)
    G, B, V0 = orc.hex_precompute(Xd)
    Amat = np.eye(3) + rng.normal(scale=0.1, size=(3, 3))
    x = Xd @ Amat.T + np.array([0.1, -0.2, 0.3])
    for s in range(8):
        F = x.T @ G[s]
        assert np.allclose(F, Amat, atol=1e-11)
        assert np.allclose(G[s].sum(0), 0, atol=1e-9)


def test_nh_small_strain_limit_is_linear_elasticity(orc):
    # Psi(I + eps H) = eps^2 (mu |sym H|^2 + lam/2 tr(H)^2) + O(eps^3): pins
    # the Lame constants (neoHookean.cpp:10-11) and the sign of every term.
    rng = np.random.default_rng(4)
    E, nu = 2.5e5, 0.37
    mu, lam = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    for _ in range(5):
        H = rng.normal(size=(3, 3))
        S = 0.5 * (H + H.T)
        lin = mu * (S * S).sum() + 0.5 * lam * np.trace(H) ** 2
        for eps in (1e-3, 1e-4):
            psi = orc.nh_psi(np.eye(3) + eps * H, mu, lam)
            assert psi / eps**2 == pytest.approx(lin, rel=50 * eps)
        # rotation invariance
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        Q *= np.sign(np.linalg.det(Q))
        F = np.eye(3) + 0.2 * H
        if np.linalg.det(F) > 0:
            assert orc.nh_psi(Q @ F, mu, lam) == pytest.approx(orc.nh_psi(F, mu, lam), rel=1e-12)


def test_nh_derivatives_finite_difference(orc):
    rng = np.random.default_rng(5)
    mu, lam = 3.4e5, 3.1e6
    F = np.eye(3) + 0.15 * rng.normal(size=(3, 3))
    P = orc.nh_dpsi(F, mu, lam).reshape(-1)
    K = orc.nh_d2psi(F, mu, lam)
    h = 1e-6
    for i in range(9):
        dF = np.zeros(9)
        dF[i] = h
        fp = orc.nh_psi((F.reshape(-1) + dF).reshape(3, 3), mu, lam)
        fm = orc.nh_psi((F.reshape(-1) - dF).reshape(3, 3), mu, lam)
        assert (fp - fm) / (2 * h) == pytest.approx(P[i], rel=1e-6, abs=1e-6 * np.abs(P).max())
        Pp = orc.nh_dpsi((F.reshape(-1) + dF).reshape(3, 3), mu, lam).reshape(-1)
        Pm = orc.nh_dpsi((F.reshape(-1) - dF).reshape(3, 3), mu, lam).reshape(-1)
        assert np.allclose((Pp - Pm) / (2 * h), K[:, i], rtol=1e-6, atol=1e-6 * np.abs(K).max())
    assert np.allclose(K, K.T, rtol=1e-13, atol=1e-13 * np.abs(K).max())


def test_inverted_element_raises(orc):
    F = np.diag([1.0, 1.0, -0.5])
    with pytest.raises(RuntimeError, match="Inverted Elements"):
        orc.nh_d2psi(F, 1.0, 1.0)


def test_project_psd_matches_eigh_clamp(orc):
    rng = np.random.default_rng(6)
    for _ in range(10):
        F = np.eye(3) + 0.4 * rng.normal(size=(3, 3))
        if np.linalg.det(F) <= 0.1:
            continue
        K = orc.nh_d2psi(F, 1.0, 8.0)
        w, V = np.linalg.eigh(K)
        ref = (V * np.maximum(w, 0)) @ V.T
        Kp = orc.project_psd(K)
        assert np.allclose(Kp, ref, atol=1e-10 * np.abs(K).max())
        assert np.linalg.eigvalsh(Kp).min() > -1e-10 * np.abs(K).max()


@pytest.mark.parametrize("kind,ed", [("tet", 4), ("hex", 8)])
def test_element_against_torch_autograd_golden(orc, kind, ed):
    d = np.load(GOLD)
    for X, x, E, nu, e, g, H in zip(d[f"{kind}_X"], d[f"{kind}_x"], d[f"{kind}_E"], d[f"{kind}_nu"],
                                    d[f"{kind}_energy"], d[f"{kind}_grad"], d[f"{kind}_hess"]):
        oe, og, oh = orc.elem_eval(X, x, E, nu)
        # at the rest state g and e vanish; measure against the stiffness
        # scale |H| h and |H| h^2 there (h = element size)
        size = np.ptp(X, axis=0).max()
        hs = np.abs(H).max()
        assert abs(oe - e) <= 1e-10 * max(abs(e), 1e-3 * hs * size**2)
        assert np.abs(og - g).max() <= 1e-10 * max(np.abs(g).max(), 1e-3 * hs * size)
        assert np.abs(oh - H).max() <= 1e-10 * hs


def test_projected_equals_raw_when_spd(orc):
    # uniform stretch with a small lambda/mu: every eigenvalue of
    # d2Psi/dF2 is positive (twist mu + (lam lnJ - mu)/s^2, flip
    # mu + (mu - lam lnJ)/s^2), so the projection must change nothing
    X = UNIT_TET * 0.01
    x = X * 1.1
    K = orc.nh_d2psi(np.eye(3) * 1.1, 1e6 / 2.2, 1e6 * 0.1 / (1.1 * 0.8))
    assert np.linalg.eigvalsh(K).min() > 0
    _, _, H = orc.elem_eval(X, x, 1e6, 0.1)
    _, _, Hp = orc.elem_eval(X, x, 1e6, 0.1, project=True)
    assert np.allclose(H, Hp, rtol=1e-12, atol=1e-12 * np.abs(H).max())
