"""Contact on the device (penalty + lagged Coulomb friction) against the
FP64 oracle: gradient, energy and Hessian at rel 1e-10 in FP64, trajectories
within 1e-6 of the bounding-box diagonal, and the reference's own
no-penetration scenario (cpp/src/_test.cpp:629-698)."""
import numpy as np
import pytest

from sors_amd import mesh, native

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def poke_pair(orc, n=6, steps=30, method="ls", **kw):
    p = mesh.contact_poke(n=n, steps=steps, **kw)
    sc = p.scene
    m_dev = native.MINIMIZE_NEWTON_LINESEARCH if method == "ls" else native.MINIMIZE_NEWTON
    m_orc = orc.NEWTON_LS if method == "ls" else orc.NEWTON
    return p, sc, m_dev, m_orc


def contact_kw(p):
    return dict(planes=[((0.0, 0.0, 0.0), (0.0, 0.0, 1.0))], disks=[(p.centers[0], (0.0, 0.0, -1.0), p.radius)],
                surface=p.surface, stiffness=p.stiffness, friction=p.friction, eps_v=p.eps_v)


def test_contact_gradient_energy_hessian_fp64(orc):
    """Start state pressed 0.2 mm into the ground (friction lag N > 0) and a
    disk already 0.3 mm inside the top face; evaluate at a perturbed q."""
    p, sc, md, mo = poke_pair(orc, n=5, gap=-3e-4)
    rng = np.random.default_rng(7)
    q0 = sc.X.reshape(-1).copy()
    q0[2::3] -= 2e-4
    r = native.Rollout.from_scene(sc, method=md)
    o = orc.System.from_scene(sc, method=mo)
    r.set_initial_deformation(q0)
    o.set_initial_deformation(q0)
    r.set_contact(**contact_kw(p))
    o.set_contact(**contact_kw(p))
    # the disk moves before the evaluated step (friction relative to it)
    c1 = p.centers[0] + np.array([2e-5, -1e-5, -1e-5])
    r.set_disk(0, c1, (0, 0, -1), p.radius)
    o.set_disk(0, c1, (0, 0, -1), p.radius)
    dt = sc.dt
    for trial in range(3):
        q = q0 + rng.normal(scale=(2e-6, 2e-5, 1e-4)[trial], size=q0.size)
        gd, go = r.gradient(q, dt), o.gradient(q, dt)
        ed, eo = r.energy(q, dt), o.energy(q, dt)
        _, _, hd = r.hessian(q, dt, native.FP64, project=True)
        _, _, ho = o.hessian(q, dt, project=True)
        say("contact parity", trial, rel_err(gd, go), abs(ed - eo) / abs(eo), rel_err(hd, ho),
            "contact energy", o.contact_energy(q, dt))
        assert o.contact_energy(q, dt) > 0
        assert rel_err(gd, go) < 1e-10
        assert abs(ed - eo) <= 1e-12 * max(abs(eo), 1.0)
        assert rel_err(hd, ho) < 1e-10


@pytest.mark.parametrize("precision,coarse", [(native.FP64, 0), (native.MIXED, 0), (native.MIXED, 32)])
def test_contact_poke_trajectory(orc, precision, coarse):
    """6^3-cell cube on the ground, disk pressing down with jitter, 30 steps:
    every state within 1e-6 of the bounding-box diagonal of the oracle
    (coarse > 0: with the two-level PCG, contact blocks included in the
    Galerkin product, refreshed every 4 tangents)."""
    p, sc, md, mo = poke_pair(orc, n=6, steps=30, gap=-1e-3, speed=0.05)
    say("poke", sc.n_nodes, sc.n_elems, "fp64" if precision == native.FP64 else "mixed", coarse)
    r = native.Rollout.from_scene(sc, method=md, precision=precision, pcg_rtol=1e-7, coarse_nodes=coarse,
                                  coarse_every=4)
    o = orc.System.from_scene(sc, method=mo)
    r.set_contact(**contact_kw(p))
    o.set_contact(**contact_kw(p))
    q = qo = sc.X.reshape(-1).copy()
    worst = 0.0
    for k in range(sc.steps):
        c = p.centers[k + 1]
        r.set_disk(0, c, (0, 0, -1), p.radius)
        o.set_disk(0, c, (0, 0, -1), p.radius)
        qo, _ = o.step(qo, sc.dt)
        q, st = r.step(q, sc.dt)
        worst = max(worst, np.abs(q - qo).max())
    say("worst |dq| / bbox", worst / sc.bbox_diag)
    x = q.reshape(-1, 3)
    c = p.centers[-1]
    under = np.linalg.norm(x[:, :2] - c[:2], axis=1) < p.radius
    # nodes under the disk stay below it (up to the penalty's give); the
    # bottom stays on the ground
    assert under.sum() > 0 and x[under, 2].max() < c[2] + 1e-4
    assert x[:, 2].min() > -1e-4
    assert worst <= 1e-6 * sc.bbox_diag


def test_reference_no_penetration_on_device():
    # cpp/src/_test.cpp:629-698 with a 1e7 N/m penalty for the hard constraint
    X = np.array([[0.0, 0.0, -0.10], [1.0, 0.0, 0.20], [0.0, 1.0, 0.20], [0.0, 0.0, 0.30]])
    conn = np.array([[0, 1, 2, 3]], np.int32)
    for precision in (native.FP64, native.MIXED):
        r = native.Rollout(X, conn, [1e5], [0.3], [1.0], grav=(0, 0, -9.81),
                           method=native.MINIMIZE_NEWTON_LINESEARCH, precision=precision, predict=0)
        r.set_contact(planes=[((0, 0, 0), (0, 0, 1))], surface=[0, 1, 2, 3], stiffness=1e7)
        q, st = r.step(X.reshape(-1).copy(), 1e-2)
        say("no-penetration z", q.reshape(-1, 3)[:, 2], st.newton_iters)
        assert q.reshape(-1, 3)[:, 2].min() >= -1e-4


def test_contact_argument_errors():
    sc = mesh.contact_poke(n=2, steps=1).scene
    r = native.Rollout.from_scene(sc)
    with pytest.raises(native.SorsError, match="out of range"):
        r.set_contact(planes=[((0, 0, 0), (0, 0, 1))], surface=[0, 10**6])
    with pytest.raises(native.SorsError, match="zero plane normal"):
        r.set_contact(planes=[((0, 0, 0), (0, 0, 0))])
    with pytest.raises(native.SorsError, match="diskIdx"):
        r.set_disk(0, (0, 0, 0), (0, 0, 1), 0.1)
    r.set_contact(planes=[((0, 0, 0), (0, 0, 1))])
    r.set_contact()  # removes contact


def test_poke_rollout_repeats_bit_for_bit():
    """Two identical contact rollouts (the bench's 40^3 poke, 500 steps:
    lagged coarse refreshes factorised on the second stream, PCG chunks
    replayed from graphs) give the same states and iteration counts bit for
    bit.  While the factorisation ran beside the solve's kernels, such runs
    parted after 79-250 steps; the solve now waits on it."""
    p = mesh.contact_poke(n=40, steps=500, seed=0)
    sc = p.scene
    cfg = dict(planes=[((0.0, 0.0, 0.0), (0.0, 0.0, 1.0))], disks=[(p.centers[0], (0.0, 0.0, -1.0), p.radius)],
               surface=p.surface, stiffness=p.stiffness, friction=p.friction, eps_v=p.eps_v)
    runs = []
    for _ in range(2):
        r = native.Rollout.from_scene(sc, precision=native.MIXED, g_tol=1e-8, pcg_rtol=3e-3, coarse_nodes=2048,
                                      coarse_every=8)
        r.set_contact(**cfg)
        its = []
        for k in range(sc.steps):
            r.set_disk(0, p.centers[k + 1], (0.0, 0.0, -1.0), p.radius)
            st = r.step_resident(sc.dt)
            its.append((st.newton_iters, st.pcg_iters))
        runs.append((r.positions(), its))
        r.close()
    print("poke repeat newton", sum(i[0] for i in runs[0][1]), sum(i[0] for i in runs[1][1]), flush=True)
    assert runs[0][1] == runs[1][1]
    assert np.array_equal(runs[0][0], runs[1][0])
