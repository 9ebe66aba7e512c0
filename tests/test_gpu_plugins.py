"""Stable Neo-Hookean and SoftCon-muscle element energies on the device
(cpp/src/energies/stableNeoHookean.cpp:3-150, softconMuscle.cpp:12-134),
through the C ABI, against the FP64 oracle (pinned in test_oracle_plugins.py).

Bars: per-element FP64 energy / gradient / Hessian to rel 1e-10 (raw and
projected); FP32 projected tangents to 1e-4 of the block scale; assembled
system gradient / energy / Hessian to rel 1e-10; trajectories within 1e-6 of
the bounding-box diagonal.
"""
import numpy as np
import pytest

from sors_amd import mesh, native

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def perturbed(sc, scale, seed=0):
    rng = np.random.default_rng(seed)
    return sc.X.reshape(-1) + rng.normal(scale=scale, size=3 * sc.n_nodes) * (1 - sc.bcmask)


def muscle_setup(sc, seed=0, groups=3):
    rng = np.random.default_rng(seed)
    E = sc.n_elems
    group = rng.integers(-1, groups, size=E).astype(np.int32)
    group[:groups] = np.arange(groups)
    stiff = np.exp(rng.uniform(np.log(1e4), np.log(3e5), size=E))
    direc = rng.normal(size=(E, 3))
    direc[::2] = [0.0, 0.0, 1.0]
    act = rng.uniform(-0.2, 0.5, size=groups)
    return group, stiff, direc, act


def oracle_elements(orc, sc, q, material, muscle, project):
    X = sc.X
    x = q.reshape(-1, 3)
    out = []
    for e in range(sc.n_elems):
        nd = sc.conn[e]
        m = None
        if muscle is not None and muscle[0][e] >= 0:
            g, k, d, a = muscle
            m = (d[e], k[e], a[g[e]])
        out.append(orc.elem_eval_ex(X[nd], x[nd], sc.youngs[e], sc.poisson[e], material, muscle=m,
                                    project=project))
    return out


def make(sc, material, muscle, **kw):
    r = native.Rollout.from_scene(sc, material=material, **kw)
    if muscle is not None:
        g, k, d, a = muscle
        r.set_muscle(g, k, d)
        r.set_muscle_activation(a)
    return r


CASES = [("snh", native.STABLE_NEOHOOKEAN, False), ("muscle", native.NEOHOOKEAN, True),
         ("snh+muscle", native.STABLE_NEOHOOKEAN, True)]


@pytest.mark.parametrize("element", ["tet", "hex"])
@pytest.mark.parametrize("name,material,with_muscle", CASES)
@pytest.mark.parametrize("project", [0, 1])
def test_element_fp64_matches_oracle(orc, element, name, material, with_muscle, project):
    sc = mesh.beam(nx=3, ny=2, nz=2, element=element, e_range=(2e5, 2e6), seed=4)
    muscle = muscle_setup(sc, seed=5) if with_muscle else None
    say("plugin element", element, name, "project", project)
    r = make(sc, material, muscle)
    q = perturbed(sc, 3e-4, seed=6)
    de, dg, dh = r.element_eval(q, precision=native.FP64, project=project)
    omat = orc.MAT_SNH if material == native.STABLE_NEOHOOKEAN else orc.MAT_NH
    ref = oracle_elements(orc, sc, q, omat, muscle, project)
    oe = np.array([t[0] for t in ref])
    og = np.array([t[1] for t in ref])
    oh = np.array([t[2] for t in ref])
    ge = np.abs(de - oe).max() / np.abs(oe).max()
    gg = np.abs(dg - og).max() / np.abs(og).max()
    gh = np.abs(dh - oh).max() / np.abs(oh).max()
    say("rel energy", ge, "grad", gg, "hess", gh)
    assert ge < 1e-10 and gg < 1e-10
    assert gh < (1e-10 if not project else 1e-9)


@pytest.mark.parametrize("element", ["tet", "hex"])
@pytest.mark.parametrize("name,material,with_muscle", CASES)
def test_element_fp32_projected_tangent(orc, element, name, material, with_muscle):
    sc = mesh.beam(nx=3, ny=2, nz=2, element=element, e_range=(2e5, 2e6), seed=7)
    muscle = muscle_setup(sc, seed=8) if with_muscle else None
    r = make(sc, material, muscle)
    q = perturbed(sc, 3e-4, seed=9)
    _, _, dh = r.element_eval(q, precision=native.MIXED, project=1)
    omat = orc.MAT_SNH if material == native.STABLE_NEOHOOKEAN else orc.MAT_NH
    oh = np.array([t[2] for t in oracle_elements(orc, sc, q, omat, muscle, True)])
    err = max(np.abs(dh[e] - oh[e]).max() / np.abs(oh[e]).max() for e in range(sc.n_elems))
    say("fp32 projected", element, name, err)
    assert err < 1e-4


@pytest.mark.parametrize("element", ["tet", "hex"])
@pytest.mark.parametrize("scheme", ["BE", "CN"])
def test_system_gradient_energy_hessian(orc, element, scheme):
    sc = mesh.beam(nx=4, ny=2, nz=2, element=element, e_range=(2e5, 2e6), seed=10)
    muscle = muscle_setup(sc, seed=11)
    sch = native.CRANK_NICOLSON if scheme == "CN" else native.BACKWARD_EULER
    r = make(sc, native.STABLE_NEOHOOKEAN, muscle, scheme=sch, method=native.MINIMIZE_NEWTON)
    o = orc.System.from_scene(sc, scheme=orc.CN if scheme == "CN" else orc.BE, method=orc.NEWTON)
    o.set_material(orc.MAT_SNH)
    g, k, d, a = muscle
    o.set_muscle(g, k, d)
    o.set_muscle_activation(a)
    q = perturbed(sc, 3e-4, seed=12)
    dt = sc.dt
    eg, og = r.gradient(q, dt), o.gradient(q, dt)
    assert np.abs(eg - og).max() <= 1e-10 * np.abs(og).max()
    ee, oe = r.energy(q, dt), o.energy(q, dt)
    assert ee == pytest.approx(oe, rel=1e-11, abs=1e-12 * abs(oe))
    rp, cols, vals = r.hessian(q, dt, precision=native.FP64, project=False)
    orp, ocols, ovals = o.hessian(q, dt)
    assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)
    assert np.abs(vals - ovals).max() <= 1e-10 * np.abs(ovals).max()
    say("system", element, scheme, "ok")


@pytest.mark.parametrize("precision", [native.FP64, native.MIXED])
def test_stable_neohookean_cantilever_trajectory(orc, precision):
    """The reference's cantilever demo material and solver
    (python/d02_cantilever.py:32-38, 60-108: stableneohookean, Crank-Nicolson,
    minimize_newton) on a clamped Kuhn beam, 40 steps."""
    sc = mesh.beam(nx=10, ny=3, nz=3, E=2.6e5, nu=0.49, rho=1070.0, steps=40)
    say("snh cantilever", precision)
    r = make(sc, native.STABLE_NEOHOOKEAN, None, scheme=native.CRANK_NICOLSON, method=native.MINIMIZE_NEWTON,
             precision=precision, pcg_rtol=1e-8)
    o = orc.System.from_scene(sc, scheme=orc.CN, method=orc.NEWTON)
    o.set_material(orc.MAT_SNH)
    q = qo = sc.X.reshape(-1).copy()
    worst = 0.0
    for _ in range(sc.steps):
        qo, _ = o.step(qo, sc.dt)
        q, _ = r.step(q, sc.dt)
        worst = max(worst, np.abs(q - qo).max())
    say("worst / bbox", worst / sc.bbox_diag)
    assert worst <= 1e-6 * sc.bbox_diag


@pytest.mark.parametrize("precision", [native.FP64, native.MIXED])
def test_muscle_actuated_hex_leg_trajectory(orc, precision):
    """Hexahedral prism clamped at its base with two SoftCon muscle groups
    along +z, driven by a two-channel activation as in
    python/d05_hopping_leg.py:136-160, 331-360 (neohookean + softconMuscle,
    backward Euler), 30 steps with minimize_newton."""
    sc = mesh.beam(nx=3, ny=3, nz=12, lx=0.03, ly=0.03, lz=0.12, element="hex", E=5e5, nu=0.4, rho=1000.0,
                   tip_speed=0.0, steps=30)
    sc.bcmask[:] = np.repeat((sc.X[:, 2] < 1e-9).astype(np.int32), 3)
    cen = sc.X[sc.conn].mean(axis=1)
    group = np.where((cen[:, 2] > 0.03) & (cen[:, 2] < 0.1), (cen[:, 0] >= 0.015).astype(np.int32), -1)
    stiff = np.full(sc.n_elems, 5e5)
    direc = np.tile([0.0, 0.0, 1.0], (sc.n_elems, 1))
    say("muscle leg", precision, int((group >= 0).sum()), "muscle elements")
    r = make(sc, native.NEOHOOKEAN, (group, stiff, direc, np.zeros(2)), method=native.MINIMIZE_NEWTON,
             precision=precision, pcg_rtol=1e-8)
    o = orc.System.from_scene(sc, method=orc.NEWTON)
    o.set_muscle(group, stiff, direc)
    q = qo = sc.X.reshape(-1).copy()
    worst = 0.0
    for k in range(sc.steps):
        act = np.array([0.3 * min(1.0, k / 10.0), 0.1 * min(1.0, k / 10.0)])
        r.set_muscle_activation(act)
        o.set_muscle_activation(act)
        qo, _ = o.step(qo, sc.dt)
        q, _ = r.step(q, sc.dt)
        worst = max(worst, np.abs(q - qo).max())
    say("worst / bbox", worst / sc.bbox_diag, "tip displacement", np.abs(qo - sc.X.reshape(-1)).max())
    assert np.abs(qo - sc.X.reshape(-1)).max() > 1e-4  # the muscles moved the leg
    assert worst <= 1e-6 * sc.bbox_diag
