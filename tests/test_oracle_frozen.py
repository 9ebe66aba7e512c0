"""Pins the oracle's minimize_frozenhessian (solver.cpp:268-294): it reaches
the same minimiser as Newton (itself pinned by the reference's known-answer
tests), and it takes more iterations on a deforming step, which a Hessian
re-assembled every iteration (plain Newton) would not."""
import numpy as np

from sors_amd import mesh


def one_step(orc, sc, method, **kw):
    s = orc.System.from_scene(sc, method=method)
    q, st = s.step(sc.X.reshape(-1), sc.dt, **kw)
    return q, st


def test_frozen_reaches_newton_minimiser(orc):
    sc = mesh.beam(nx=8, ny=2, nz=2, element="tet", e_range=(5e5, 5e6), seed=3)
    tight = dict(dx_tol=1e-13)
    qn, sn = one_step(orc, sc, orc.NEWTON, **tight)
    qf, sf = one_step(orc, sc, orc.FROZEN, **tight)
    err = np.abs(qf - qn).max() / sc.bbox_diag
    print("frozen vs newton", err, sf.newton_iters, sn.newton_iters, flush=True)
    assert err < 1e-10
    assert np.abs(qn - sc.X.reshape(-1)).max() > 1e-6  # the step moved
    assert sf.newton_iters > sn.newton_iters


def test_frozen_without_convergence_returns_iterate(orc):
    # two iterations cannot reach dx_tol: the iterate comes back, no error
    sc = mesh.beam(nx=6, ny=2, nz=2, element="tet", e_range=(5e5, 5e6), seed=3)
    q, st = one_step(orc, sc, orc.FROZEN, max_iter=2, dx_tol=1e-15)
    assert st.newton_iters == 2 and np.all(np.isfinite(q))
