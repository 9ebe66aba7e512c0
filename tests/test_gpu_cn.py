"""Crank-Nicolson on the device (cpp/src/integrator.cpp:27-36, 70-79,
108-121; integrator.h:79-97) against the FP64 oracle, whose CN branch is
pinned by closed forms in test_oracle_cn.py.

The reference's own CN callers (python/d01_msd_timestepping.py:24-25,
python/d02_cantilever.py:37-38) use minimize_newton, as these tests do.
"""
import numpy as np
import pytest

from sors_amd import mesh, native

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("element", ["tet", "hex"])
@pytest.mark.parametrize("precision", [native.FP64, native.MIXED])
def test_cn_trajectory_matches_oracle(orc, element, precision):
    """Clamped beam under gravity with the tip-velocity profile, 40 CN steps:
    every state within 1e-6 of the bounding-box diagonal of the oracle's
    (LDL^T on the raw Hessian)."""
    sc = mesh.beam(nx=10, ny=3, nz=3, element=element, E=5e5, steps=40)
    say("cn trajectory", element, "fp64" if precision == native.FP64 else "mixed")
    o = orc.System.from_scene(sc, scheme=orc.CN, method=orc.NEWTON)
    r = native.Rollout.from_scene(sc, scheme=native.CRANK_NICOLSON, method=native.MINIMIZE_NEWTON,
                                  precision=precision, pcg_rtol=1e-8)
    q = qo = sc.X.reshape(-1).copy()
    worst = 0.0
    for _ in range(sc.steps):
        qo, _ = o.step(qo, sc.dt)
        q, _ = r.step(q, sc.dt)
        worst = max(worst, np.abs(q - qo).max())
    qs, vs = r.state()
    qos, vos = o.state()
    say("worst |dq| / bbox", worst / sc.bbox_diag, "|dv|", np.abs(vs - vos).max())
    assert worst <= 1e-6 * sc.bbox_diag
    # the CN velocity 2 (q - q_prev) / dt - v_prev carries the position
    # difference amplified by 2 / dt, and does not damp it: 1e-5 of |v|
    assert np.abs(vs - vos).max() <= 1e-5 * np.abs(vos).max()


def test_cn_free_fall_closed_form_on_device():
    """No strain, constant acceleration: the trapezoidal rule is exact."""
    sc = mesh.beam(nx=4, ny=2, nz=2)
    sc.bcmask[:] = 0
    v0 = np.tile([0.1, -0.05, 0.3], sc.n_nodes)
    sc.v0 = v0
    r = native.Rollout.from_scene(sc, scheme=native.CRANK_NICOLSON, method=native.MINIMIZE_NEWTON,
                                  precision=native.FP64, pcg_rtol=1e-12)
    dt, steps = 1e-3, 12
    q = sc.X.reshape(-1).copy()
    for _ in range(steps):
        q, _ = r.step(q, dt)
    t = steps * dt
    exact = sc.X.reshape(-1) + v0 * t + 0.5 * np.tile(sc.grav, sc.n_nodes) * t * t
    say("cn free fall |q - exact|", np.abs(q - exact).max())
    assert np.abs(q - exact).max() < 1e-11


def test_cn_small_vibration_keeps_its_energy_on_device(orc):
    """Linear regime: kinetic + strain energy of the device's CN states stays
    within 1e-6 (the oracle's restatement holds it to about 1e-8; backward
    Euler loses 16% in the same 60 steps)."""
    sc = mesh.beam(nx=6, ny=2, nz=2, E=1e5)
    sc.bcmask[:] = 0
    sc.grav = np.zeros(3)
    x = sc.X[:, 0] - sc.X[:, 0].mean()
    v0 = np.zeros((sc.n_nodes, 3))
    v0[:, 2] = 1e-3 * (x ** 2 / (x ** 2).mean() - 1.0)
    sc.v0 = v0.reshape(-1)
    r = native.Rollout.from_scene(sc, scheme=native.CRANK_NICOLSON, method=native.MINIMIZE_NEWTON,
                                  precision=native.FP64, pcg_rtol=1e-12)
    o = orc.System.from_scene(sc, scheme=orc.CN, method=orc.NEWTON)
    m = o.mass()

    def mech(q, v):
        return 0.5 * (m * v * v).sum() + r.energy(q, 0.0)

    q = sc.X.reshape(-1).copy()
    e0 = mech(q, sc.v0)
    worst = 0.0
    for _ in range(60):
        q, _ = r.step(q, 2e-4)
        qq, v = r.state()
        worst = max(worst, abs(mech(qq, v) - e0) / e0)
    say("cn energy drift", worst)
    assert worst < 1e-6
