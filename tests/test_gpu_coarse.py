"""Two-level PCG preconditioner (options.coarse_nodes > 0): block-Jacobi
plus a coarse correction on the rigid-body modes of spatial aggregates.

The preconditioner changes the Krylov iterates, never the system solved:
solutions match a dense direct solve and trajectories match the FP64 oracle
to the same bars as block-Jacobi, with fewer PCG iterations.
"""
import numpy as np
import pytest

from sors_amd import mesh, native

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("precision,element", [(native.FP64, "tet"), (native.MIXED, "tet"), (native.FP64, "hex"),
                                               (native.MIXED, "hex")])
def test_coarse_linear_solve_matches_direct(orc, precision, element):
    sc = mesh.beam(nx=16, ny=3, nz=3, element=element, e_range=(5e5, 5e6), seed=1)
    o = orc.System.from_scene(sc, method=orc.NEWTON)
    rng = np.random.default_rng(4)
    q = sc.X.reshape(-1) + rng.normal(scale=1e-4, size=3 * sc.n_nodes) * (1 - sc.bcmask)
    dt = 1e-2
    H = o.hessian_dense(q, dt, project=True)
    b = rng.normal(size=3 * sc.n_nodes) * (1 - sc.bcmask)
    xr = np.linalg.solve(H, b)
    rtol = 1e-10 if precision == native.FP64 else 1e-6
    its = {}
    for cn in (0, 16):
        r = native.Rollout.from_scene(sc, method=native.MINIMIZE_NEWTON, precision=precision, coarse_nodes=cn)
        x, it = r.linear_solve(q, dt, b, rtol=rtol)
        err = np.abs(x - xr).max() / np.abs(xr).max()
        say("coarse_nodes", cn, "pcg iters", it, "rel err", err)
        assert err < (1e-8 if precision == native.FP64 else 1e-4)
        its[cn] = it
        r.close()
    assert its[16] < 0.8 * its[0]


@pytest.mark.parametrize("cn", [300, 100000])
def test_coarse_aggregates_wider_than_a_workgroup(cn):
    """Aggregates of several hundred nodes (the restriction sums them over
    several wavefronts) and the coarsest space (target above the node count:
    three cubic bins along the beam): the FP64 residual |H x - b| of the
    solve meets the requested tolerance, in fewer iterations than
    block-Jacobi."""
    sp = pytest.importorskip("scipy.sparse")
    sc = mesh.beam(nx=40, ny=6, nz=6, e_range=(5e5, 5e6), seed=5)
    rng = np.random.default_rng(6)
    q = sc.X.reshape(-1) + rng.normal(scale=1e-5, size=3 * sc.n_nodes) * (1 - sc.bcmask)
    b = rng.normal(size=3 * sc.n_nodes) * (1 - sc.bcmask)
    n = 3 * sc.n_nodes
    its = {}
    for k in (0, cn):
        r = native.Rollout.from_scene(sc, method=native.MINIMIZE_NEWTON, precision=native.FP64, coarse_nodes=k)
        x, its[k] = r.linear_solve(q, 1e-3, b, rtol=1e-10)
        rp, ci, v = r.hessian(q, 1e-3, native.FP64, project=True)
        H = sp.csr_matrix((v, ci, rp), shape=(n, n))
        res = np.linalg.norm(H @ x - b) / np.linalg.norm(b)
        say("coarse_nodes", k, "iters", its[k], "relative residual", res)
        assert res <= 2e-10
        r.close()
    assert its[cn] < its[0]


@pytest.mark.parametrize("precision,every", [(native.FP64, 1), (native.MIXED, 1), (native.MIXED, 3)])
def test_coarse_cantilever_trajectory(orc, precision, every):
    """20x4x4 Kuhn beam, 200 steps: every state within 1e-6 of the
    bounding-box diagonal of the FP64 oracle; fewer PCG iterations in total
    than block-Jacobi alone.  every = 3: the coarse operator is refreshed on
    every third tangent, factorised on the second stream ahead of a solve
    that runs with the previous inverse."""
    sc = mesh.cantilever_small()
    o = orc.System.from_scene(sc, method=orc.NEWTON_LS)
    r = native.Rollout.from_scene(sc, precision=precision, pcg_rtol=1e-6, coarse_nodes=24, coarse_every=every)
    b = native.Rollout.from_scene(sc, precision=precision, pcg_rtol=1e-6)
    q = qo = qb = sc.X.reshape(-1).copy()
    worst = 0.0
    it_c = it_b = 0
    for _ in range(sc.steps):
        qo, _ = o.step(qo, sc.dt)
        q, st = r.step(q, sc.dt)
        qb, sb = b.step(qb, sc.dt)
        it_c += st.pcg_iters
        it_b += sb.pcg_iters
        worst = max(worst, np.abs(q - qo).max())
    say("coarse trajectory worst/bbox", worst / sc.bbox_diag, "pcg iters", it_c, "block-Jacobi", it_b)
    assert worst <= 1e-6 * sc.bbox_diag
    assert it_c < it_b


@pytest.mark.parametrize("precision,every", [(native.FP64, 1), (native.MIXED, 1), (native.MIXED, 2)])
def test_coarse_graph_replay_matches_direct_launches(precision, every):
    """The coarse kernels inside captured PCG chunks: graph replays and
    direct launches agree bit for bit (every = 2: with lagged refreshes
    factorised on the second stream as well)."""
    sc = mesh.beam(nx=12, ny=3, nz=3, e_range=(5e5, 5e6), seed=2)
    kw = dict(pcg_check_every=16, precision=precision, coarse_nodes=16, coarse_every=every)
    a = native.Rollout.from_scene(sc, use_graph=0, **kw)
    b = native.Rollout.from_scene(sc, use_graph=1, **kw)
    qa = qb = sc.X.reshape(-1).copy()
    for _ in range(6):
        qa, sa = a.step(qa, sc.dt)
        qb, sb = b.step(qb, sc.dt)
        say("coarse graph replay", precision, every, sa.pcg_iters, sb.pcg_iters)
        assert sa.pcg_iters == sb.pcg_iters
        assert np.array_equal(qa, qb)


def test_coarse_option_switches_per_rollout():
    """Changing coarse_nodes between steps rebuilds (or drops) the coarse
    space; each setting reproduces a fresh rollout with that setting."""
    sc = mesh.beam(nx=10, ny=3, nz=3, e_range=(5e5, 5e6), seed=3)
    q0 = sc.X.reshape(-1).copy()
    ref = {}
    for cn in (0, 8, 32):
        r = native.Rollout.from_scene(sc, coarse_nodes=cn)
        ref[cn] = r.step(q0, sc.dt)[0]
        r.close()
    r = native.Rollout.from_scene(sc, coarse_nodes=8)
    v0 = sc.v0.copy()
    for cn in (8, 32, 0, 8):
        r.set_options(coarse_nodes=cn)
        r.set_state(q0, v0)  # the velocity and the predictor's history restart
        q, _ = r.step(q0, sc.dt)
        say("coarse_nodes switched to", cn, "max |dq|", np.abs(q - ref[cn]).max())
        assert np.array_equal(q, ref[cn])


@pytest.mark.parametrize("nx,n,cn", [(16, 3, 64), (16, 3, 16), (40, 8, 16)])
def test_coarse_inverse_matches_dense_inverse(nx, n, cn):
    """The hand-written blocked Gauss-Jordan inverse of the coarse operator E
    (padded to the 64-row tile: one tile and 2 pivot steps, 4 x 4 tiles, and
    39 x 39 tiles with 78 steps) against numpy's inverse of the same matrix
    (E read from its lower triangle): exactly symmetric, and within a
    condition-scaled bound of the dense inverse and of E Einv = I."""
    sc = mesh.beam(nx=nx, ny=n, nz=n, e_range=(5e5, 5e6), seed=2)
    r = native.Rollout.from_scene(sc, precision=native.MIXED, coarse_nodes=cn, coarse_every=1)
    r.step(sc.X.reshape(-1).copy(), sc.dt)
    E, Einv = r.coarse_operator()
    r.close()
    nc = E.shape[0]
    Es = np.tril(E) + np.tril(E, -1).T
    ref = np.linalg.inv(Es)
    cond = np.linalg.cond(Es)
    err = np.abs(Einv - ref).max() / np.abs(ref).max()
    res = np.abs(Es @ Einv - np.eye(nc)).max()
    say("nc", nc, "cond", cond, "inverse rel err", err, "residual", res)
    assert np.array_equal(Einv, Einv.T)
    assert err <= 1e-13 * cond + 1e-14
    assert res <= 1e-13 * cond + 1e-14


def test_fully_clamped_mesh_steps_with_coarse_option():
    """Every node Dirichlet-fixed: the coarse space has no aggregates, the
    solve falls back to block-Jacobi and the step keeps every node at its
    prescribed position (the reference steps such a mesh trivially)."""
    sc = mesh.beam(nx=6, ny=2, nz=2)
    sc.bcmask = np.ones_like(sc.bcmask)
    r = native.Rollout.from_scene(sc, precision=native.MIXED, coarse_nodes=2048)
    q, st = r.step(sc.X.reshape(-1).copy(), sc.dt)
    assert r.coarse_operator() is None
    r.close()
    assert np.array_equal(q, sc.X.reshape(-1))
