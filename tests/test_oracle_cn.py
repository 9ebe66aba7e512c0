"""Pins the oracle's Crank-Nicolson branch (cpp/src/integrator.cpp:27-36,
70-79, 108-121; integrator.h:79-97) by what the mathematics fixes:

- the trapezoidal rule is exact for constant acceleration, so a body in free
  fall (no strain) follows q0 + v0 t + g t^2 / 2 to rounding, where backward
  Euler is off by g dt t / 2;
- with no stiffness, the CN energy's kinetic (dt^2 / 2) and damping (1/4)
  factors are the ones whose derivative is the CN gradient (central
  differences);
- for a linear elastic system the trapezoidal rule conserves kinetic plus
  strain energy exactly, so the 1/2 (g_int + g_prev) split of the internal
  force keeps a small-amplitude vibration's energy to about 1e-8 while
  backward Euler loses 16% of it in 60 steps;
- damping enters at the new velocity (first order), so the damped free fall
  converges to v(t) = g/a + (v0 - g/a) e^{-a t} with error O(dt).
"""
import numpy as np
import pytest

from sors_amd import mesh


def free_beam(nx=4, ny=2, nz=2, E=1e6, grav=(0.0, 0.0, -9.81), alpha=0.0):
    sc = mesh.beam(nx=nx, ny=ny, nz=nz, E=E)
    sc.bcmask[:] = 0
    sc.grav = np.array(grav, float)
    sc.alpha = alpha
    return sc


def test_cn_free_fall_is_exact_for_constant_acceleration(orc):
    sc = free_beam()
    v0 = np.tile([0.1, -0.05, 0.3], sc.n_nodes)
    dt, steps = 1e-3, 12
    out = {}
    for scheme in (orc.CN, orc.BE):
        s = orc.System(sc.X, sc.conn, sc.youngs, sc.poisson, sc.density, grav=sc.grav, scheme=scheme,
                       method=orc.NEWTON)
        s.set_initial_velocity(v0)
        q = sc.X.reshape(-1).copy()
        for _ in range(steps):
            q, _ = s.step(q, dt)
        out[scheme] = q
    t = steps * dt
    exact = sc.X.reshape(-1) + v0 * t + 0.5 * np.tile(sc.grav, sc.n_nodes) * t * t
    assert np.abs(out[orc.CN] - exact).max() < 1e-12
    # backward Euler lags by g dt t / 2 in z
    assert np.abs(out[orc.BE] - exact).max() == pytest.approx(0.5 * 9.81 * dt * t, rel=1e-6)


def test_cn_energy_derivative_is_cn_gradient_without_stiffness(orc):
    sc = free_beam(E=0.0, alpha=0.7)
    s = orc.System(sc.X, sc.conn, sc.youngs, sc.poisson, sc.density, grav=sc.grav, alpha=0.7, scheme=orc.CN,
                   method=orc.NEWTON)
    rng = np.random.default_rng(3)
    s.set_initial_deformation(sc.X.reshape(-1) + 1e-3 * rng.standard_normal(3 * sc.n_nodes))
    s.set_initial_velocity(0.2 * rng.standard_normal(3 * sc.n_nodes))
    q = sc.X.reshape(-1) + 1e-3 * rng.standard_normal(3 * sc.n_nodes)
    dt = 2e-3
    g = s.gradient(q, dt)
    h = 1e-7
    for j in range(0, 3 * sc.n_nodes, 7):
        e = np.zeros_like(q)
        e[j] = h
        fd = (s.energy(q + e, dt) - s.energy(q - e, dt)) / (2 * h)
        assert fd == pytest.approx(g[j], rel=1e-6, abs=1e-9 * np.abs(g).max())


def _mech_energy(s, q, v):
    # kinetic + strain energy (energy at dt = 0 is strain + gravity; no gravity here)
    return 0.5 * (s.mass() * v * v).sum() + s.energy(q, 0.0)


def test_cn_conserves_linear_elastic_energy_where_be_dissipates(orc):
    sc = free_beam(nx=6, ny=2, nz=2, E=1e5, grav=(0.0, 0.0, 0.0))
    # a small bending velocity with no net momentum: linear regime (strain ~1e-5)
    x = sc.X[:, 0] - sc.X[:, 0].mean()
    v0 = np.zeros((sc.n_nodes, 3))
    v0[:, 2] = 1e-3 * (x ** 2 / (x ** 2).mean() - 1.0)
    dt, steps = 2e-4, 60
    drift = {}
    for scheme in (orc.CN, orc.BE):
        s = orc.System(sc.X, sc.conn, sc.youngs, sc.poisson, sc.density, grav=sc.grav, scheme=scheme,
                       method=orc.NEWTON)
        s.set_initial_velocity(v0.reshape(-1))
        q = sc.X.reshape(-1).copy()
        e0 = _mech_energy(s, q, v0.reshape(-1))
        worst = 0.0
        for _ in range(steps):
            q, _ = s.step(q, dt)
            qq, v = s.state()
            worst = max(worst, abs(_mech_energy(s, qq, v) - e0) / e0)
        drift[scheme] = worst
    assert drift[orc.CN] < 1e-6
    assert drift[orc.BE] > 0.05


def test_cn_damped_free_fall_converges_at_first_order(orc):
    a, g0, vz0, T = 3.0, -9.81, 0.4, 0.05
    sc = free_beam(nx=2, ny=1, nz=1, E=1e5, alpha=a)
    exact = g0 / a + (vz0 - g0 / a) * np.exp(-a * T)
    errs = []
    for dt in (2e-3, 1e-3):
        s = orc.System(sc.X, sc.conn, sc.youngs, sc.poisson, sc.density, grav=sc.grav, alpha=a, scheme=orc.CN,
                       method=orc.NEWTON)
        s.set_initial_velocity(np.tile([0.0, 0.0, vz0], sc.n_nodes))
        q = sc.X.reshape(-1).copy()
        for _ in range(int(round(T / dt))):
            q, _ = s.step(q, dt)
        vz = s.state()[1].reshape(-1, 3)[:, 2]
        assert np.ptp(vz) < 1e-12
        errs.append(abs(vz[0] - exact))
    assert errs[1] < 0.6 * errs[0] and errs[1] > 0.4 * errs[0]
    assert errs[1] < 1e-2 * abs(exact)
