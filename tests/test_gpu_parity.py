"""Parity of the HIP path (through the C ABI) with the CPU FP64 oracle.

Bars: CSR pattern and scatter map bit-exact; per-element FP64 energy /
gradient / Hessian to rel 1e-10; FP32 element tangents to rel 1e-4 of the
block scale; converged positions within 1e-6 of the bounding-box diagonal.
"""
import os

import numpy as np
import pytest

from sors_amd import mesh, native

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "element_nh.npz")


def say(*a):
    print(*a, flush=True)


def rollout(sc, **kw):
    return native.Rollout.from_scene(sc, **kw)


def perturbed(sc, scale, seed=0):
    rng = np.random.default_rng(seed)
    q = sc.X.reshape(-1) + rng.normal(scale=scale, size=3 * sc.n_nodes) * (1 - sc.bcmask)
    return q


def rel_err(a, b, scale=None):
    s = np.abs(b).max() if scale is None else scale
    return np.abs(a - b).max() / max(s, 1e-300)


@pytest.mark.parametrize("element", ["tet", "hex"])
def test_element_fp64_matches_oracle(orc, element):
    sc = mesh.beam(nx=4, ny=2, nz=3, element=element, e_range=(5e5, 5e6), seed=3)
    say("element_eval", element, sc.n_elems)
    r = rollout(sc)
    q = perturbed(sc, 3e-4)
    e, g, H = r.element_eval(q, native.FP64, project=False)
    ed = sc.ed
    for k in range(sc.n_elems):
        Xe = sc.X[sc.conn[k]]
        xe = q.reshape(-1, 3)[sc.conn[k]]
        oe, og, oh = orc.elem_eval(Xe, xe, sc.youngs[k], sc.poisson[k])
        hs = np.abs(oh).max()
        size = np.ptp(Xe, axis=0).max()
        assert abs(e[k] - oe) <= 1e-10 * max(abs(oe), 1e-3 * hs * size**2)
        assert np.abs(g[k] - og).max() <= 1e-10 * max(np.abs(og).max(), 1e-3 * hs * size)
        assert np.abs(H[k] - oh).max() <= 1e-10 * hs
    assert g.shape == (sc.n_elems, 3 * ed)


@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_element_golden_vectors_fp64(kind):
    d = np.load(GOLD)
    Xs, xs, Es, nus = d[f"{kind}_X"], d[f"{kind}_x"], d[f"{kind}_E"], d[f"{kind}_nu"]
    ed = Xs.shape[1]
    for X, x, E, nu, e, g, H in zip(Xs, xs, Es, nus, d[f"{kind}_energy"], d[f"{kind}_grad"], d[f"{kind}_hess"]):
        conn = np.arange(ed, dtype=np.int32)[None]
        r = native.Rollout(X, conn, [E], [nu], [1000.0])
        de, dg, dh = r.element_eval(x.reshape(-1), native.FP64)
        hs = np.abs(H).max()
        size = np.ptp(X, axis=0).max()
        assert abs(de[0] - e) <= 1e-10 * max(abs(e), 1e-3 * hs * size**2)
        assert np.abs(dg[0] - g).max() <= 1e-10 * max(np.abs(g).max(), 1e-3 * hs * size)
        assert np.abs(dh[0] - H).max() <= 1e-10 * hs
        r.close()


@pytest.mark.parametrize("element", ["tet", "hex"])
def test_projected_tangent_matches_eigen_clamp(orc, element):
    sc = mesh.beam(nx=3, ny=2, nz=2, element=element, E=3e5, nu=0.45)
    r = rollout(sc)
    q = perturbed(sc, 8e-4, seed=1)
    _, _, H64 = r.element_eval(q, native.FP64, project=True)
    _, _, H32 = r.element_eval(q, native.MIXED, project=True)
    worst32 = 0.0
    for k in range(sc.n_elems):
        Xe = sc.X[sc.conn[k]]
        xe = q.reshape(-1, 3)[sc.conn[k]]
        _, _, oh = orc.elem_eval(Xe, xe, sc.youngs[k], sc.poisson[k], project=True)
        hs = np.abs(oh).max()
        assert np.abs(H64[k] - oh).max() <= 1e-9 * hs
        worst32 = max(worst32, np.abs(H32[k] - oh).max() / hs)
        assert np.linalg.eigvalsh(0.5 * (H64[k] + H64[k].T)).min() >= -1e-9 * hs
    say("projected FP32 tangent worst rel err", worst32)
    assert worst32 < 1e-4


@pytest.mark.parametrize("case", ["tet_clamp", "hex_clamp", "tet_partial", "tet_sqp", "tet_dt0"])
def test_pattern_and_scatter_map_bit_exact(orc, case):
    element = "hex" if case.startswith("hex") else "tet"
    sc = mesh.beam(nx=4, ny=2, nz=3, element=element)
    mask = sc.bcmask.copy()
    if case == "tet_partial":
        mask[:] = 0
        mask[[0, 7, 8, 20, 33]] = 1  # per-DOF mask: the general (full) map path
    method = native.MINIMIZE_SQP if case == "tet_sqp" else native.MINIMIZE_NEWTON
    omethod = orc.SQP if case == "tet_sqp" else orc.NEWTON
    dt = 0.0 if case == "tet_dt0" else 1e-3
    say("pattern", case)
    r = native.Rollout(sc.X, sc.conn, sc.youngs, sc.poisson, sc.density, mask, sc.bcvalue, sc.grav, method=method)
    o = orc.System(sc.X, sc.conn, sc.youngs, sc.poisson, sc.density, mask, sc.bcvalue, sc.grav, method=omethod)
    rp, cols = r.pattern(dt)
    orp, ocols = o.pattern(dt)
    assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)
    assert np.array_equal(r.scatter_map(dt), o.scatter_map(dt))


def test_pattern_and_map_cantilever_small(orc):
    sc = mesh.cantilever_small()
    r = rollout(sc)
    o = orc.System.from_scene(sc)
    rp, cols = r.pattern(sc.dt)
    orp, ocols = o.pattern(sc.dt)
    assert len(cols) == 52437
    assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)
    assert np.array_equal(r.scatter_map(sc.dt), o.scatter_map(sc.dt))


@pytest.mark.parametrize("element", ["tet", "hex"])
@pytest.mark.parametrize("scheme", ["be", "cn"])
def test_gradient_energy_hessian_fp64(orc, element, scheme):
    sc = mesh.beam(nx=4, ny=2, nz=2, element=element, e_range=(5e5, 5e6), seed=5)
    osch = orc.BE if scheme == "be" else orc.CN
    dsch = native.BACKWARD_EULER if scheme == "be" else native.CRANK_NICOLSON
    r = rollout(sc, scheme=dsch, method=native.MINIMIZE_NEWTON)
    o = orc.System.from_scene(sc, scheme=osch, method=orc.NEWTON)
    q = perturbed(sc, 2e-4, seed=2)
    dt = 1e-3
    gd, go = r.gradient(q, dt), o.gradient(q, dt)
    assert rel_err(gd, go) < 1e-10
    ed_, eo = r.energy(q, dt), o.energy(q, dt)
    assert abs(ed_ - eo) <= 1e-12 * max(abs(eo), 1.0)
    _, _, hd = r.hessian(q, dt, native.FP64, project=False)
    _, _, ho = o.hessian(q, dt)
    assert rel_err(hd, ho) < 1e-10
    _, _, hm = r.hessian(q, dt, native.MIXED, project=False)
    say("FP32 Hessian rel err", rel_err(hm, ho))
    assert rel_err(hm, ho) < 1e-5


def test_reference_dummy_energy_tests_on_device():
    # cpp/src/_test.cpp:136-176 with the polynomial test element on the device
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0, -1, 0]], float)
    conn = np.array([[0, 1, 2, 3], [0, 1, 4, 3]], np.int32)
    r = native.Rollout(X, conn, [0.0, 0.0], [0.0, 0.0], [1.0, 1.0], scheme=native.CRANK_NICOLSON,
                       method=native.MINIMIZE_NEWTON, material=native.TEST_POLY)
    q = X.reshape(-1)
    assert r.energy(q, 0.0) == 3.0
    gradGT = np.array([[0, 0, 0], [6, 0, 0], [0, 3, 0], [0, 0, 6], [0, -3, 0]], float).reshape(-1)
    assert np.allclose(r.gradient(q, 0.0), gradGT)
    hess = np.array([[-2, -2, -2], [22, -2, -2], [-1, 11, -1], [-2, -2, 22], [-1, 11, -1]], float).reshape(-1)
    assert np.allclose(r.hessian_dense(q, 0.0), np.diag(hess))


def test_cube_free_fall_on_device():
    # cpp/src/_test.cpp:395-443
    X = np.array([[0, 0, 0], [0, 1, 0], [1, 1, 0], [1, 0, 0], [0, 0, 1], [0, 1, 1], [1, 1, 1], [1, 0, 1]], float)
    conn = np.array([[0, 1, 3, 4], [1, 4, 5, 6], [1, 2, 3, 6], [3, 4, 6, 7], [1, 3, 4, 6]], np.int32)
    for prec in (native.FP64, native.MIXED):
        r = native.Rollout(X, conn, np.full(5, 5e5), np.full(5, 0.4), np.ones(5), grav=(0, 0, -10),
                           method=native.MINIMIZE_NEWTON, precision=prec)
        q = X.reshape(-1).copy()
        com = [q.reshape(-1, 3).mean(0)[2]]
        for _ in range(9):
            q, st = r.step(q, 0.01)
            com.append(q.reshape(-1, 3).mean(0)[2])
        com = np.array(com)
        acc = (-2 * com[1:-1] + com[2:] + com[:-2]) / 0.01**2
        assert acc[1:6].mean() == pytest.approx(-10.0, abs=1e-3)


@pytest.mark.parametrize("precision", [native.FP64, native.MIXED])
def test_linear_solve_matches_direct(orc, precision):
    sc = mesh.beam(nx=6, ny=2, nz=2, E=1e6)
    r = rollout(sc, method=native.MINIMIZE_NEWTON, precision=precision)
    o = orc.System.from_scene(sc, method=orc.NEWTON)
    q = perturbed(sc, 1e-4, seed=4)
    dt = 1e-3
    H = o.hessian_dense(q, dt, project=True)
    rng = np.random.default_rng(9)
    b = rng.normal(size=3 * sc.n_nodes)
    x, it = r.linear_solve(q, dt, b, rtol=1e-10 if precision == native.FP64 else 1e-6)
    xr = np.linalg.solve(H, b)
    err = np.abs(x - xr).max() / np.abs(xr).max()
    say("pcg iters", it, "rel err", err)
    assert err < (1e-8 if precision == native.FP64 else 1e-4)


@pytest.mark.parametrize("precision", [native.FP64, native.MIXED])
def test_cantilever_small_trajectory(orc, precision):
    """20x4x4 Kuhn beam, 200 implicit steps, Newton + Armijo: every state
    within 1e-6 of the bounding-box diagonal of the FP64 oracle (LDL^T on the
    raw Hessian, the reference's algorithm)."""
    sc = mesh.cantilever_small()
    say("cantilever_small", "fp64" if precision == native.FP64 else "mixed")
    o = orc.System.from_scene(sc, method=orc.NEWTON_LS)
    r = rollout(sc, precision=precision, pcg_rtol=1e-6)
    q = qo = sc.X.reshape(-1).copy()
    worst = 0.0
    for k in range(sc.steps):
        qo, _ = o.step(qo, sc.dt)
        q, st = r.step(q, sc.dt)
        worst = max(worst, np.abs(q - qo).max())
    say("worst |dq| / bbox diag", worst / sc.bbox_diag)
    assert worst <= 1e-6 * sc.bbox_diag


def test_cantilever_small_trajectory_bench_solver(orc):
    """The same 200 steps in mixed precision with the bench's solver settings
    (PCG rtol 3e-3, 3e-2 for each step's first Newton solve, a solve aimed at
    ||g|| < 1e-8 N reusing the last tangent, two-level preconditioner): still
    within 1e-6 of the bounding-box diagonal of the oracle; the looser first
    solves take fewer PCG iterations than rtol 3e-3 throughout, and the reuse
    assembles fewer tangents."""
    sc = mesh.cantilever_small()
    o = orc.System.from_scene(sc, method=orc.NEWTON_LS)
    base = dict(precision=native.MIXED, pcg_rtol=3e-3, g_tol=1e-8, coarse_nodes=64, coarse_every=4,
                profile_kernels=1)
    cases = {"bench": dict(pcg_rtol_first=3e-2, reuse_tangent=1), "no_first": dict(reuse_tangent=1),
             "no_reuse": dict(pcg_rtol_first=3e-2)}
    runs = {k: rollout(sc, **base, **v) for k, v in cases.items()}
    q = {k: sc.X.reshape(-1).copy() for k in runs}
    qo = sc.X.reshape(-1).copy()
    worst = {k: 0.0 for k in runs}
    pcg = {k: 0 for k in runs}
    tangents = {k: 0 for k in runs}
    for _ in range(sc.steps):
        qo, _ = o.step(qo, sc.dt)
        for k, r in runs.items():
            q[k], st = r.step(q[k], sc.dt)
            pcg[k] += st.pcg_iters
            tangents[k] += st.assemble_launches
            worst[k] = max(worst[k], np.abs(q[k] - qo).max())
    say("worst |dq| / bbox diag", {k: w / sc.bbox_diag for k, w in worst.items()}, "PCG", pcg, "tangents", tangents)
    assert max(worst.values()) <= 1e-6 * sc.bbox_diag
    assert pcg["bench"] < pcg["no_first"]
    assert tangents["bench"] < tangents["no_reuse"]


@pytest.mark.parametrize("precision", [native.FP64, native.MIXED])
def test_graph_replay_matches_direct_launches(precision):
    """PCG chunks replayed from a captured hipGraph give the same states as
    direct launches, across several replays with host work in between.  Tet
    assembly on the sliced node-block layout is atomic-free and every
    reduction has a fixed order, so the two runs agree bit for bit."""
    sc = mesh.beam(nx=12, ny=3, nz=3, e_range=(5e5, 5e6), seed=2)
    a = rollout(sc, use_graph=0, pcg_check_every=16, precision=precision)
    b = rollout(sc, use_graph=1, pcg_check_every=16, precision=precision)
    qa = qb = sc.X.reshape(-1).copy()
    for _ in range(4):
        qa, sa = a.step(qa, sc.dt)
        qb, sb = b.step(qb, sc.dt)
        say("graph replay", precision, sa.pcg_iters, sb.pcg_iters, np.abs(qa - qb).max() / sc.bbox_diag)
        assert sa.pcg_iters == sb.pcg_iters
        assert np.array_equal(qa, qb)


@pytest.mark.parametrize("element", ["tet", "hex"])
def test_residual_check_after_aimed_solve_is_gradient_only(element):
    """After a solve whose forcing term came from g_tol, Newton checks the
    next residual with a gradient-only assembly and assembles the tangent
    only if the check fails.  With profiling on, the timed (tangent)
    assemblies then number one per Newton solve, not one more per step, and
    the residual that ended each step is the true one: ||g|| of an FP64
    gradient evaluation at the returned state, on a second rollout that
    still holds the step's history (and then takes the same step, bit for
    bit)."""
    sc = mesh.beam(nx=12, ny=3, nz=3, element=element, e_range=(5e5, 5e6), seed=2)
    # g_tol such that steps end on the residual criterion (||g|| ~ 1e-7 N
    # when the step-size criterion would stop them)
    g_tol = 1e-6
    opts = dict(g_tol=g_tol, pcg_rtol=3e-3)
    r = rollout(sc, profile_kernels=1, **opts)
    lag = rollout(sc, **opts)
    q = sc.X.reshape(-1).copy()
    solves = tangents = ended_on_g = 0
    for _ in range(4):
        q_new, st = r.step(q, sc.dt)
        solves += st.newton_iters
        tangents += st.assemble_launches
        say("step", element, st.newton_iters, st.assemble_launches, st.gnorm)
        assert st.newton_iters <= st.assemble_launches <= st.newton_iters + 1
        if st.gnorm < g_tol:
            ended_on_g += 1
            g = lag.gradient(q_new, sc.dt)
            assert abs(np.linalg.norm(g) - st.gnorm) <= 1e-12 * st.gnorm
        q_lag, _ = lag.step(q, sc.dt)
        assert np.array_equal(q_lag, q_new)
        q = q_new
    assert ended_on_g > 0
    assert tangents < solves + 4  # at least one final check skipped the tangent


@pytest.mark.parametrize("precision", [native.FP64, native.MIXED])
@pytest.mark.parametrize("element", ["tet", "hex"])
def test_frozen_hessian_trajectory(orc, element, precision):
    """minimize_frozenhessian (solver.cpp:268-294): one tangent per step,
    gradient-only iterations; states within 1e-6 of the bounding-box
    diagonal of the oracle's frozen iteration, and more iterations than
    Newton takes on the same steps."""
    sc = mesh.beam(nx=8, ny=2, nz=2, element=element, e_range=(5e5, 5e6), seed=3, steps=10)
    o = orc.System.from_scene(sc, method=orc.FROZEN)
    r = rollout(sc, method=native.MINIMIZE_FROZENHESSIAN, precision=precision, pcg_rtol=1e-8)
    rn = rollout(sc, method=native.MINIMIZE_NEWTON, precision=precision, pcg_rtol=1e-8)
    q = qn = qo = sc.X.reshape(-1).copy()
    worst, it_f, it_n = 0.0, 0, 0
    for _ in range(sc.steps):
        qo, _ = o.step(qo, sc.dt)
        q, st = r.step(q, sc.dt)
        qn, stn = rn.step(qn, sc.dt)
        it_f += st.newton_iters
        it_n += stn.newton_iters
        worst = max(worst, np.abs(q - qo).max())
    say("frozen", element, precision, worst / sc.bbox_diag, it_f, it_n)
    assert worst <= 1e-6 * sc.bbox_diag
    assert it_f > it_n
    assert np.abs(q - sc.X.reshape(-1)).max() > 1e-6


def test_mixed_solve_that_does_not_converge_goes_on_in_fp64():
    """A mixed-precision Newton solve that reaches max_newton without
    meeting the tolerances continues from its iterate with FP64 tangents and
    solves, and ends at the FP64 solution."""
    sc = mesh.beam(nx=8, ny=2, nz=2, element="tet", e_range=(5e5, 5e6), seed=3)
    tol = dict(g_tol=0.0, dx_tol=1e-10, pcg_rtol=1e-8)
    r = rollout(sc, precision=native.MIXED, max_newton=3, **tol)
    r64 = rollout(sc, precision=native.FP64, **tol)
    q0 = sc.X.reshape(-1).copy()
    q, st = r.step(q0, sc.dt)
    q64, st64 = r64.step(q0, sc.dt)
    say("fallback", st.fp64_fallbacks, st.newton_iters, st64.newton_iters, np.abs(q - q64).max() / sc.bbox_diag)
    assert st.fp64_fallbacks == 1 and st64.fp64_fallbacks == 0
    assert st.newton_iters > 3
    assert np.abs(q - q64).max() <= 1e-9 * sc.bbox_diag


@pytest.mark.parametrize("precision", [native.FP64, native.MIXED])
def test_kernel_timing_does_not_change_results(precision):
    """profile_kernels launches the first PCG chunk of every solve and the
    assembly kernels with timing events (created with a device-scope
    release); the states must equal an untimed run bit for bit."""
    sc = mesh.beam(nx=30, ny=6, nz=6, e_range=(5e5, 5e6), seed=6)
    runs = []
    for prof in (0, 1):
        r = rollout(sc, precision=precision, profile_kernels=prof, g_tol=1e-8, pcg_rtol=3e-3)
        q = sc.X.reshape(-1).copy()
        iters = []
        for _ in range(3):
            q, st = r.step(q, sc.dt)
            iters.append(st.pcg_iters)
        runs.append((q, iters))
        r.close()
    say("timed vs untimed", precision, runs[0][1], runs[1][1])
    assert runs[0][1] == runs[1][1]
    assert np.array_equal(runs[0][0], runs[1][0])


@pytest.mark.parametrize("element", ["tet", "hex"])
def test_node_assembly_is_deterministic(element):
    """Two assemblies of the same state give identical gradients and
    Hessian values (no float atomics on the node-block path: element rows,
    then fixed-order sums per layout slot)."""
    sc = mesh.beam(nx=8, ny=3, nz=3, element=element, e_range=(5e5, 5e6), seed=4)
    r = rollout(sc)
    q = perturbed(sc, 3e-4, seed=5)
    for prec in (native.FP64, native.MIXED):
        _, _, h1 = r.hessian(q, sc.dt, prec, project=True)
        _, _, h2 = r.hessian(q, sc.dt, prec, project=True)
        assert np.array_equal(h1, h2)
    assert np.array_equal(r.gradient(q, sc.dt), r.gradient(q, sc.dt))


@pytest.mark.parametrize("precision", [native.FP64, native.MIXED])
def test_per_dof_mask_csr_path_trajectory(orc, precision):
    """A mask that is not node-granular keeps the scalar CSR layout (general
    scatter map, CSR SpMV); the trajectory still matches the oracle."""
    sc = mesh.beam(nx=8, ny=2, nz=2, steps=10)
    mask = sc.bcmask.copy()
    clamp = np.nonzero(sc.bcmask)[0]
    mask[clamp[2::3]] = 0  # x=0 nodes keep z free: the clamp holds x and y only
    r = native.Rollout(sc.X, sc.conn, sc.youngs, sc.poisson, sc.density, mask, sc.bcvalue, sc.grav,
                       precision=precision, pcg_rtol=1e-8)
    r.set_initial_velocity(sc.v0)
    o = orc.System(sc.X, sc.conn, sc.youngs, sc.poisson, sc.density, mask, sc.bcvalue, sc.grav,
                   method=orc.NEWTON_LS)
    o.set_initial_velocity(sc.v0)
    q = qo = sc.X.reshape(-1).copy()
    for _ in range(sc.steps):
        q, _ = r.step(q, sc.dt)
        qo, _ = o.step(qo, sc.dt)
    assert np.abs(q - qo).max() <= 1e-6 * sc.bbox_diag


def test_block_layout_export_matches_csr_assembly(orc):
    """The sliced node-block values exported to CSR order equal the oracle's
    assembled Hessian (projected), for the node-granular clamp."""
    sc = mesh.beam(nx=5, ny=2, nz=2, element="hex", E=4e5)
    r = rollout(sc, method=native.MINIMIZE_NEWTON)
    o = orc.System.from_scene(sc, method=orc.NEWTON)
    q = perturbed(sc, 5e-4, seed=8)
    _, _, hd = r.hessian(q, sc.dt, native.FP64, project=True)
    _, _, ho = o.hessian(q, sc.dt, project=True)
    assert rel_err(hd, ho) < 1e-9


def test_forcing_options_are_validated():
    sc = mesh.beam(nx=2, ny=1, nz=1)
    r = rollout(sc)
    for bad in (dict(pcg_rtol_first=1.5), dict(pcg_rtol_first=-0.1), dict(reuse_tangent=2)):
        with pytest.raises(native.SorsError):
            r.set_options(**bad)
    r.set_options(pcg_rtol_first=3e-2, reuse_tangent=1)
    o = r.options()
    assert o.pcg_rtol_first == 3e-2 and o.reuse_tangent == 1
    r.close()


def test_spmv_column_offsets_and_their_fallback():
    """The SpMV reads 16-bit column offsets when the node ordering allows
    (a structured beam) and the 32-bit table when it does not (the same beam
    with its nodes shuffled: offsets beyond 32,767); the isolated SpMV's
    algorithmic bytes show which, and both orderings step to the same state."""
    sc = mesh.beam(nx=100, ny=18, nz=18, e_range=(5e5, 5e6), seed=2, steps=1)
    N = sc.n_nodes
    assert N > 32768
    perm = np.random.default_rng(0).permutation(N)      # new index of each node
    inv = np.argsort(perm)
    X = sc.X.reshape(-1, 3)
    mask = np.asarray(sc.bcmask).reshape(-1, 3)
    v0 = np.asarray(sc.v0).reshape(-1, 3)
    runs = {}
    for name, order in (("structured", np.arange(N)), ("shuffled", inv)):
        where = np.argsort(order)                      # position of node k in the new numbering
        r = native.Rollout(X[order], where[sc.conn], sc.youngs, sc.poisson, sc.density,
                           mask[order].reshape(-1), np.zeros(3 * N), sc.grav, precision=native.FP64,
                           pcg_rtol=1e-10, g_tol=1e-10)
        r.set_initial_velocity(v0[order].reshape(-1))
        q, _ = r.step(X[order].reshape(-1), sc.dt)
        by = r.bench_kernel(native.KERNEL_SPMV, reps=2)[1]
        runs[name] = (by, q.reshape(-1, 3)[where])
        r.close()
    (by_s, q_s), (by_p, q_p) = runs["structured"], runs["shuffled"]
    nblocks = (by_p - by_s) / 2.0          # 4 B against 2 B per stored block
    say("SpMV bytes", by_s, by_p, "blocks", nblocks, "max |dq|/bbox", np.abs(q_s - q_p).max() / sc.bbox_diag)
    assert nblocks > N and nblocks == int(nblocks)
    assert np.abs(q_s - q_p).max() <= 1e-8 * sc.bbox_diag
