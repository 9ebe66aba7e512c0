"""Pins the oracle's Stable Neo-Hookean and SoftCon-muscle restatements
(cpp/src/energies/stableNeoHookean.cpp:3-150, softconMuscle.cpp:12-134)
against torch FP64 autograd golden vectors (tests/golden/make_golden_plugins.py)
and what the mathematics fixes:

- Stable Neo-Hookean: energy / gradient / Hessian equal autograd of the
  closed-form density; its reparameterisation makes the rest state stress
  free and its rest tangent the linear-elastic tensor of the original Lame
  constants (mu (d_ip d_jq + d_iq d_jp) + lam d_ij d_pq);
- SoftCon muscle: the reference's energy is k/2 |l - L|, while its gradient
  is that of k/2 (l - L)^2 (autograd of the squared form), and its Hessian
  equals the exact Hessian of that squared form plus
  k V0 sum_s w (1 - r_s) (t_s t_s^T - sum_a du_a du_a^T), where du_a = d(F m)_a/dx
  and t_s = sum_a du_a: the reference sums the rows of dF/dx before the
  outer product (softconMuscle.cpp:108-115).
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "element_plugins.npz")


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_snh_matches_autograd_golden(orc, kind):
    d = np.load(GOLD)
    for c in range(len(d[f"{kind}_E"])):
        X, x, E, nu = d[f"{kind}_X"][c], d[f"{kind}_x"][c], d[f"{kind}_E"][c], d[f"{kind}_nu"][c]
        e, g, H = orc.elem_eval_ex(X, x, E, nu, orc.MAT_SNH)
        ge, gg, gh = d[f"{kind}_snh_energy"][c], d[f"{kind}_snh_grad"][c], d[f"{kind}_snh_hess"][c]
        # at the rest state g vanishes: measure against the stiffness scale |H| h
        size = np.ptp(X, axis=0).max()
        hs = np.abs(gh).max()
        assert e == pytest.approx(ge, rel=1e-11)
        assert np.abs(g - gg).max() <= 1e-10 * max(np.abs(gg).max(), 1e-3 * hs * size)
        assert rel(H, gh) < 1e-10


def test_snh_rest_state_and_linear_limit(orc):
    E, nu = 3e5, 0.42
    mu, lam = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    I = np.eye(3)
    assert np.abs(orc.snh_dpsi(I, mu, lam)).max() < 1e-9 * mu
    K = orc.snh_d2psi(I, mu, lam)
    d = np.eye(3)
    C = (mu * (np.einsum("ip,jq->ijpq", d, d) + np.einsum("iq,jp->ijpq", d, d)) +
         lam * np.einsum("ij,pq->ijpq", d, d)).reshape(9, 9)
    assert np.allclose(K, C, rtol=1e-12, atol=1e-9 * lam)
    # Neo-Hookean has the same rest tangent
    assert np.allclose(orc.nh_d2psi(I, mu, lam), C, rtol=1e-12, atol=1e-9 * lam)
    # stays finite through inversion (no log J)
    F = np.diag([1.0, 1.0, -0.2])
    assert np.isfinite(orc.snh_psi(F, mu, lam)) and np.isfinite(orc.snh_d2psi(F, mu, lam)).all()


def _muscle_rows(orc, kind, X, x, m):
    """Per Gauss point: weight, r-independent du_a/dx (3 x nd) and F m."""
    if kind == "tet":
        _, V0, B = orc.tet_precompute(X)
        Bs, w = [B], 1.0
    else:
        G, B8, V0 = orc.hex_precompute(X)
        Bs, w = list(B8), 1.0 / 8
    out = []
    xe = x.reshape(-1)
    for B in Bs:
        F = (B @ xe).reshape(3, 3)
        du = np.stack([sum(B[3 * a + c] * m[c] for c in range(3)) for a in range(3)])
        out.append((w, du, F @ m))
    return V0, out


@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_muscle_matches_autograd_golden_and_reference_hessian_form(orc, kind):
    d = np.load(GOLD)
    for c in range(len(d[f"{kind}_E"])):
        X, x = d[f"{kind}_X"][c], d[f"{kind}_x"][c]
        m, k, act = d[f"{kind}_m"][c], d[f"{kind}_k"][c], d[f"{kind}_act"][c]
        # E = 0: the elastic term vanishes, the element carries the muscle alone
        e, g, H = orc.elem_eval_ex(X, x, 0.0, 0.3, orc.MAT_NH, muscle=(m, k, act))
        assert e == pytest.approx(d[f"{kind}_muscle_energy"][c], rel=1e-11)
        assert rel(g, d[f"{kind}_muscle_sq_grad"][c]) < 1e-10
        V0, pts = _muscle_rows(orc, kind, X, x, m)
        corr = 0.0
        for w, du, u in pts:
            r = (1 - act) / np.linalg.norm(u)
            t = du.sum(0)
            corr = corr + k * V0 * w * (1 - r) * (np.outer(t, t) - du.T @ du)
        assert rel(H, d[f"{kind}_muscle_sq_hess"][c] + corr) < 1e-10
        # the projected Hessian is PSD and equals the reference's where that is PSD
        _, _, Hp = orc.elem_eval_ex(X, x, 0.0, 0.3, orc.MAT_NH, muscle=(m, k, act), project=True)
        assert np.linalg.eigvalsh(Hp).min() > -1e-9 * np.abs(Hp).max()
        if np.linalg.eigvalsh(H).min() > 1e-9 * np.abs(H).max():
            assert rel(Hp, H) < 1e-9


def test_snh_projected_hessian_is_psd_with_rigid_null_space(orc):
    d = np.load(GOLD)
    for c in range(len(d["tet_E"])):
        X, x, E, nu = d["tet_X"][c], d["tet_x"][c], d["tet_E"][c], d["tet_nu"][c]
        _, _, H = orc.elem_eval_ex(X, x, E, nu, orc.MAT_SNH)
        _, _, Hp = orc.elem_eval_ex(X, x, E, nu, orc.MAT_SNH, project=True)
        ev = np.linalg.eigvalsh(Hp)
        assert ev.min() > -1e-9 * ev.max()
        # rigid translations stay in the null space
        t = np.tile([1.0, 0.0, 0.0], 4)
        assert np.abs(Hp @ t).max() < 1e-8 * np.abs(Hp).max()
