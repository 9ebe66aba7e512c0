"""Writes tests/golden/element_nh.npz: per-element Neo-Hookean energy,
gradient and Hessian for tet4 and hex8, from torch FP64 autograd.

The quantities are defined independently of the oracle's code:
  * tet4: F = Ds Dm^-1 with Ds/Dm the edge matrices against node 3
    (the definition in cpp/src/geometries/tetrahedron.cpp:152-169);
  * hex8: F at the 8 Gauss points (+-1/sqrt(3)) of the isoparametric
    trilinear map, F = (dx/dxi)(dX/dxi)^-1, with dx/dxi taken by autograd of
    x(xi) = sum_i N_i(xi) x_i, N_i = (1 + xi_i xi)(1 + eta_i eta)(1 + zeta_i zeta)/8
    and the reference's corner signs (cpp/src/geometries/hexahedron.cpp:171-179);
    element energy = V0 * mean_s Psi(F_s) with V0 from the reference's 5-tet
    split (hexahedron.cpp:106-130, neoHookean.cpp:26-37);
  * Psi = mu/2 (tr F^T F - 3) - mu ln J + lam/2 (ln J)^2 with the Lame
    constants of neoHookean.cpp:10-11.
Gradient and Hessian come from torch.autograd, so they check the oracle's
hand-derived dPsi/dF and d2Psi/dF2 contractions.

Run:  python tests/golden/make_golden.py
"""
import os

import numpy as np
import torch

torch.set_default_dtype(torch.float64)
HERE = os.path.dirname(os.path.abspath(__file__))

XI = torch.tensor([[1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, -1],
                   [1, -1, 1], [1, 1, 1], [-1, 1, 1], [-1, -1, 1]], dtype=torch.float64)
HEX_TETS = [[0, 1, 3, 4], [1, 2, 3, 6], [1, 4, 5, 6], [3, 4, 6, 7], [1, 3, 4, 6]]


def lame(E, nu):
    return E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))


def psi(F, mu, lam):
    J = torch.det(F)
    lj = torch.log(J)
    return 0.5 * mu * ((F * F).sum() - 3) - mu * lj + 0.5 * lam * lj * lj


def tet_energy(X, x, E, nu):
    mu, lam = lame(E, nu)
    Dm = torch.stack([X[0] - X[3], X[1] - X[3], X[2] - X[3]], dim=1)
    Ds = torch.stack([x[0] - x[3], x[1] - x[3], x[2] - x[3]], dim=1)
    F = Ds @ torch.linalg.inv(Dm)
    V0 = torch.abs(torch.dot(X[1] - X[0], torch.linalg.cross(X[2] - X[0], X[3] - X[0]))) / 6
    return V0 * psi(F, mu, lam)


def trilinear(nodes, xi):
    N = 0.125 * (1 + XI[:, 0] * xi[0]) * (1 + XI[:, 1] * xi[1]) * (1 + XI[:, 2] * xi[2])
    return (N[:, None] * nodes).sum(0)


def jac(nodes, xi):
    return torch.autograd.functional.jacobian(lambda z: trilinear(nodes, z), xi, create_graph=True)


def hex_energy(X, x, E, nu):
    mu, lam = lame(E, nu)
    g = 1 / np.sqrt(3)
    V0 = 0
    for t in HEX_TETS:
        a, b, c, d = (X[i] for i in t)
        V0 = V0 + torch.abs(torch.dot(b - a, torch.linalg.cross(c - a, d - a))) / 6
    e = 0
    for s in range(8):
        xi = g * XI[s].clone()
        F = jac(x, xi) @ torch.linalg.inv(jac(X, xi))
        e = e + psi(F, mu, lam) / 8
    return V0 * e


def cases(kind, rng, count):
    out = []
    for c in range(count):
        if kind == "tet":
            X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) * rng.uniform(0.003, 0.02)
            X = X + rng.normal(scale=0.002, size=X.shape) * (c > 0)
        else:
            h = rng.uniform(0.002, 0.01, size=3)
            X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                          [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float) * h
            X = X + rng.normal(scale=0.05 * h.min(), size=X.shape) * (c > 0)
        scale = [0.0, 0.02, 0.1, 0.25][c % 4]
        A = np.eye(3) + rng.normal(scale=scale, size=(3, 3))
        if np.linalg.det(A) <= 0.2:
            A = np.eye(3) + 0.05 * rng.normal(size=(3, 3))
        x = X @ A.T + rng.normal(scale=1e-3, size=3) + rng.normal(scale=scale * 0.05 * X.std(), size=X.shape)
        E = float(np.exp(rng.uniform(np.log(1e5), np.log(1e7))))
        nu = float(rng.uniform(0.3, 0.48))
        out.append((X, x, E, nu))
    return out


def main():
    rng = np.random.default_rng(20260101)
    data = {}
    for kind, fn, count in [("tet", tet_energy, 12), ("hex", hex_energy, 6)]:
        Xs, xs, Es, nus, en, gr, he = [], [], [], [], [], [], []
        for X, x, E, nu in cases(kind, rng, count):
            Xt = torch.tensor(X)
            xt = torch.tensor(x).reshape(-1)

            def f(v):
                return fn(Xt, v.reshape(-1, 3), E, nu)

            e = f(xt)
            g = torch.autograd.functional.jacobian(f, xt)
            H = torch.autograd.functional.hessian(f, xt)
            Xs.append(X)
            xs.append(x)
            Es.append(E)
            nus.append(nu)
            en.append(float(e))
            gr.append(g.numpy())
            he.append(H.numpy())
        data[f"{kind}_X"] = np.array(Xs)
        data[f"{kind}_x"] = np.array(xs)
        data[f"{kind}_E"] = np.array(Es)
        data[f"{kind}_nu"] = np.array(nus)
        data[f"{kind}_energy"] = np.array(en)
        data[f"{kind}_grad"] = np.array(gr)
        data[f"{kind}_hess"] = np.array(he)
    np.savez_compressed(os.path.join(HERE, "element_nh.npz"), **data)
    print("wrote", os.path.join(HERE, "element_nh.npz"))


if __name__ == "__main__":
    main()
