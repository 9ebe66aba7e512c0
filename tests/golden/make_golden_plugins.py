"""Writes tests/golden/element_plugins.npz: per-element Stable Neo-Hookean and
SoftCon-muscle quantities for tet4 and hex8, from torch FP64 autograd.

Geometry, F and V0 are defined as in make_golden.py (tet4 edge matrices;
hex8 isoparametric map with autograd Jacobians at the 8 Gauss points, V0 from
the 5-tet split).  The energy densities are the reference's closed forms:

  * Stable Neo-Hookean (cpp/src/energies/stableNeoHookean.cpp:9-14, 31-40):
        mu' = 4/3 mu, lam' = lam + 5/8 mu', alpha = 1 + 3/4 mu'/lam',
        Psi = mu'/2 (I_C - 3) + lam'/2 (J - alpha)^2 - mu'/2 log(I_C + 1);
  * SoftCon muscle (cpp/src/energies/softconMuscle.cpp:27-31) with direction
    m, stiffness k, activation e, l = |F m|, L = 1 - e:
        the reference's energy   k/2 |l - L|            ("muscle_energy")
        and the potential of its gradient k/2 (l - L)^2  ("muscle_sq_*"):
    autograd of the second gives the gradient the reference computes
    (:55-72) and the exact Hessian of that potential; the reference's own
    Hessian (:98-130) differs from it by a closed-form term that the test
    states.

Run:  python tests/golden/make_golden_plugins.py
"""
import os

import numpy as np
import torch

from make_golden import HEX_TETS, cases, jac, lame

torch.set_default_dtype(torch.float64)
HERE = os.path.dirname(os.path.abspath(__file__))


def snh_psi(F, mu, lam):
    mus = 4.0 / 3.0 * mu
    lams = lam + 5.0 / 8.0 * mus
    alpha = 1 + 0.75 * mus / lams
    IC = (F * F).sum()
    J = torch.det(F)
    return mus / 2 * (IC - 3) + lams / 2 * (J - alpha) ** 2 - mus / 2 * torch.log(IC + 1)


def muscle_abs(F, m, k, e):
    l = torch.linalg.norm(F @ m)
    return k / 2 * torch.abs(l - (1 - e))


def muscle_sq(F, m, k, e):
    l = torch.linalg.norm(F @ m)
    return k / 2 * (l - (1 - e)) ** 2


def tet_F(X, x):
    Dm = torch.stack([X[0] - X[3], X[1] - X[3], X[2] - X[3]], dim=1)
    Ds = torch.stack([x[0] - x[3], x[1] - x[3], x[2] - x[3]], dim=1)
    V0 = torch.abs(torch.dot(X[1] - X[0], torch.linalg.cross(X[2] - X[0], X[3] - X[0]))) / 6
    return [Ds @ torch.linalg.inv(Dm)], V0


def hex_F(X, x):
    g = 1 / np.sqrt(3)
    from make_golden import XI
    V0 = 0
    for t in HEX_TETS:
        a, b, c, d = (X[i] for i in t)
        V0 = V0 + torch.abs(torch.dot(b - a, torch.linalg.cross(c - a, d - a))) / 6
    Fs = []
    for s in range(8):
        xi = g * XI[s].clone()
        Fs.append(jac(x, xi) @ torch.linalg.inv(jac(X, xi)))
    return Fs, V0


def element(kind, X, x, density):
    Fs, V0 = (tet_F if kind == "tet" else hex_F)(X, x)
    return V0 * sum(density(F) for F in Fs) / len(Fs)


def main():
    rng = np.random.default_rng(20261204)
    data = {}
    for kind, count in [("tet", 12), ("hex", 6)]:
        rows = {k: [] for k in ("X", "x", "E", "nu", "m", "k", "act", "snh_energy", "snh_grad", "snh_hess",
                                "muscle_energy", "muscle_sq_grad", "muscle_sq_hess")}
        for c, (X, x, E, nu) in enumerate(cases(kind, rng, count)):
            mu, lam = lame(E, nu)
            m = rng.normal(size=3)
            m = m / np.linalg.norm(m) if c % 3 else np.array([0.0, 0.0, 1.0])
            k = float(np.exp(rng.uniform(np.log(1e4), np.log(1e6))))
            # activations up to 0.6; every third case strongly activated, so
            # that l > L on some and l < L (r > 1) on others
            act = float(rng.uniform(0.0, 0.6)) if c % 3 != 2 else float(rng.uniform(-0.4, -0.1))
            Xt = torch.tensor(X)
            xt = torch.tensor(x).reshape(-1)
            mt = torch.tensor(m)

            def fs(v):
                return element(kind, Xt, v.reshape(-1, 3), lambda F: snh_psi(F, mu, lam))

            def fa(v):
                return element(kind, Xt, v.reshape(-1, 3), lambda F: muscle_abs(F, mt, k, act))

            def fq(v):
                return element(kind, Xt, v.reshape(-1, 3), lambda F: muscle_sq(F, mt, k, act))

            rows["X"].append(X)
            rows["x"].append(x)
            rows["E"].append(E)
            rows["nu"].append(nu)
            rows["m"].append(m)
            rows["k"].append(k)
            rows["act"].append(act)
            rows["snh_energy"].append(float(fs(xt)))
            rows["snh_grad"].append(torch.autograd.functional.jacobian(fs, xt).numpy())
            rows["snh_hess"].append(torch.autograd.functional.hessian(fs, xt).numpy())
            rows["muscle_energy"].append(float(fa(xt)))
            rows["muscle_sq_grad"].append(torch.autograd.functional.jacobian(fq, xt).numpy())
            rows["muscle_sq_hess"].append(torch.autograd.functional.hessian(fq, xt).numpy())
        for key, val in rows.items():
            data[f"{kind}_{key}"] = np.array(val)
    out = os.path.join(HERE, "element_plugins.npz")
    np.savez_compressed(out, **data)
    print("wrote", out)


if __name__ == "__main__":
    main()
