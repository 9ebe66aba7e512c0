"""Serving loop (continuous batching) through the C ABI.

Each rollout's result must be the same alone and in the mix, against a plain
device-resident rollout stepped on its own.  Tet assembly on the node-block
layout is atomic-free and a batched PCG launch keeps each rollout's own
block counts and reduction order, so the two agree bit for bit.
"""
import numpy as np
import pytest

from sors_amd import mesh, native

pytestmark = pytest.mark.gpu

SOLVER = dict(g_tol=1e-8, pcg_rtol=3e-3)


def say(*a):
    print(*a, flush=True)


def small_request(i, seed=11):
    return mesh.serving_request(i, seed=seed, n_lo=300, n_hi=3000, steps_lo=3, steps_hi=6)


def alone(sc, precision=native.MIXED, **extra):
    r = native.Rollout.from_scene(sc, precision=precision, **dict(SOLVER, **extra))
    for _ in range(sc.steps):
        r.step_resident(sc.dt)
    q = r.positions().reshape(-1, 3)
    r.close()
    return q


def test_serving_mix_matches_alone():
    reqs = [small_request(i) for i in range(8)]
    say("serving mix", [(sc.name, sc.n_nodes, sc.steps) for sc in reqs])
    srv = native.Server(max_active=5, prep_threads=3, pcg_chunk=8, **SOLVER)
    ids = [srv.submit(sc, sc.steps, sc.dt) for sc in reqs]
    out = [srv.wait(i, timeout=120.0) for i in ids]
    st = srv.stats()
    srv.close()
    say("server stats", st)
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert 1 <= st["peak_active"] <= 5
    assert st["batched_launches"] > 0
    for sc, (q, rs) in zip(reqs, out):
        ref = alone(sc)
        err = np.abs(q - ref).max() / sc.bbox_diag
        say(sc.name, "mix vs alone", err, rs["newton_iters"], rs["pcg_iters"], rs["latency_ms"])
        assert rs["steps"] == sc.steps and rs["status"] == native.OK
        assert np.array_equal(q, ref)
        assert np.abs(q - sc.X).max() > 1e-6  # the rollout moved


@pytest.mark.parametrize("every", [1, 3])
def test_serving_two_level_pcg_matches_alone(every):
    """Served rollouts with the coarse correction (each rollout its own
    aggregates; batched restriction and apply launches) equal stand-alone
    rollouts with the same options bit for bit, and take fewer PCG
    iterations than block-Jacobi alone."""
    coarse = dict(coarse_nodes=64, coarse_every=every)
    reqs = [small_request(i, seed=7) for i in range(6)]
    srv = native.Server(max_active=4, prep_threads=2, pcg_chunk=8, **SOLVER, **coarse)
    ids = [srv.submit(sc, sc.steps, sc.dt) for sc in reqs]
    out = [srv.wait(i, timeout=120.0) for i in ids]
    srv.close()
    bj = native.Server(max_active=4, prep_threads=2, pcg_chunk=8, coarse_nodes=0, **SOLVER)
    ids = [bj.submit(sc, sc.steps, sc.dt) for sc in reqs]
    out_bj = [bj.wait(i, timeout=120.0) for i in ids]
    bj.close()
    it_c = it_b = 0
    for sc, (q, rs), (_, rb) in zip(reqs, out, out_bj):
        ref = alone(sc, **coarse)
        say(sc.name, sc.n_nodes, "coarse served vs alone", np.abs(q - ref).max(), rs["pcg_iters"], rb["pcg_iters"])
        assert rs["status"] == native.OK
        assert np.array_equal(q, ref)
        it_c += rs["pcg_iters"]
        it_b += rb["pcg_iters"]
    assert it_c < it_b


def test_serving_bench_solver_options_match_alone():
    """pcg_rtol_first (a looser first Newton solve per step) and reuse_tangent
    (a solve aimed at g_tol may reuse the last tangent) are applied by the
    served Newton phases exactly as by sors_step: served equals alone bit for
    bit, and the looser first solves take fewer PCG iterations in total."""
    first = dict(pcg_rtol_first=3e-2, reuse_tangent=1, coarse_nodes=64, coarse_every=2)
    reqs = [small_request(i, seed=5) for i in range(6)]
    res = {}
    for name, extra in (("first", first), ("plain", dict(first, pcg_rtol_first=0.0))):
        srv = native.Server(max_active=4, prep_threads=2, pcg_chunk=8, **SOLVER, **extra)
        ids = [srv.submit(sc, sc.steps, sc.dt) for sc in reqs]
        res[name] = [srv.wait(i, timeout=120.0) for i in ids]
        srv.close()
    it = {name: sum(rs["pcg_iters"] for _, rs in out) for name, out in res.items()}
    say("PCG iterations", it)
    for sc, (q, rs) in zip(reqs, res["first"]):
        assert rs["status"] == native.OK
        assert np.array_equal(q, alone(sc, **first))
    assert it["first"] < it["plain"]


def test_serving_single_request_matches_plain_fp64():
    sc = small_request(3)
    srv = native.Server(max_active=2, prep_threads=1, precision=native.FP64, **SOLVER)
    q, rs = srv.wait(srv.submit(sc, sc.steps, sc.dt), timeout=120.0)
    srv.close()
    ref = alone(sc, precision=native.FP64)
    err = np.abs(q - ref).max() / sc.bbox_diag
    say("fp64 server vs alone", err)
    assert np.array_equal(q, ref)


def test_serving_budget_limits_residency():
    reqs = [small_request(i, seed=5) for i in range(6)]
    # budget for about two of the largest requests at once: not all fit
    est = [native_estimate(sc) for sc in reqs]
    per = max(est)
    assert sum(est) > 2.5 * per
    srv = native.Server(hbm_budget_bytes=2.5 * per, max_active=8, prep_threads=2, coarse_nodes=0, **SOLVER)
    ids = [srv.submit(sc, sc.steps, sc.dt) for sc in reqs]
    for i in ids:
        srv.wait(i, timeout=120.0)
    st = srv.stats()
    srv.close()
    say("budget stats", st, "budget", 2.5 * per)
    assert st["completed"] == len(reqs)
    assert 0 < st["peak_resident_bytes"] <= 2.5 * per
    assert st["peak_active"] < len(reqs)
    assert st["resident_bytes"] == 0.0  # everything released


def test_serving_reports_failures():
    sc = small_request(0)
    srv = native.Server(max_active=2, prep_threads=1, max_newton=1, g_tol=1e-14, pcg_rtol=1e-6)
    rid = srv.submit(sc, sc.steps, sc.dt)
    with pytest.raises(native.SorsError) as ei:
        srv.wait(rid, timeout=120.0)
    assert ei.value.code == native.ERR_NOCONV
    assert "did not converge" in str(ei.value)
    # the server keeps serving after a failed request
    srv2 = native.Server(max_active=2, prep_threads=1, **SOLVER)
    q, rs = srv2.wait(srv2.submit(sc, 2, sc.dt), timeout=120.0)
    assert rs["steps"] == 2
    srv.close()
    srv2.close()


def native_estimate(sc):
    """Device bytes the server reserves for a request before its pattern exists
    (block-Jacobi PCG: coarse_nodes = 0)."""
    n, N, E, ed = 3 * sc.n_nodes, sc.n_nodes, sc.n_elems, sc.ed
    nblk = N * (24 if ed == 4 else 32)
    # two-phase assembly: element rows and slot lists
    rows = E * (4 * ed * 8 + 10 * ed * ed * 8 + ed * ed * 4) + 4 * nblk
    return (n * (19 * 8 + 4) + 72 * N + nblk * (9 * 8 + 4 + 9 * 4) + E * (ed * 4 + 16 + ed * ed) + rows
            + 8 * 9 * 2048 + (1 << 20))


def test_served_stream_requests_match_cpu_oracle(orc):
    """Requests 0, 2 and 5 of the seeded serving stream (seed 0; the stream's
    generator with node counts drawn in [2k, 20k]: a 3,321-node beam, an
    8,000-node cube and a 3,375-node cube, each with its own E and nu),
    served together in mixed precision and checked against the FP64 CPU
    oracle after 4 steps each: within 1e-6 of the bounding-box diagonal."""
    reqs = []
    for i in (0, 2, 5):
        sc = mesh.serving_request(i, seed=0, n_lo=2000, n_hi=20000)
        sc.steps = 4
        reqs.append(sc)
    say("served vs oracle", [(sc.name, sc.n_nodes) for sc in reqs])
    srv = native.Server(max_active=4, prep_threads=3, pcg_chunk=16, **SOLVER)
    ids = [srv.submit(sc, sc.steps, sc.dt) for sc in reqs]
    out = [srv.wait(i, timeout=300.0) for i in ids]
    srv.close()
    for sc, (q, rs) in zip(reqs, out):
        o = orc.System.from_scene(sc, method=orc.NEWTON_LS, nthreads=16)
        qo = sc.X.reshape(-1).copy()
        for _ in range(sc.steps):
            qo, _ = o.step(qo, sc.dt, linear_solver=orc.PCG, project=1, pcg_rtol=1e-9, g_tol=1e-9)
        err = np.abs(q.reshape(-1) - qo).max() / sc.bbox_diag
        say(sc.name, sc.n_nodes, "served vs oracle", err, "moved", np.abs(qo - sc.X.reshape(-1)).max())
        assert rs["status"] == native.OK and rs["steps"] == sc.steps
        assert err <= 1e-6


def test_serving_fp64_continuation_matches_alone():
    """A mixed solve that reaches max_newton goes on from its iterate in FP64,
    served as alone (sors_step), so the served result equals the stand-alone
    one bit for bit and the request reports the fallback."""
    # loose inner solves converge linearly (about 20x per Newton iteration):
    # four mixed iterations leave the residual above g_tol, four more in FP64
    # from that iterate bring it below
    opts = dict(g_tol=1e-9, pcg_rtol=5e-2, max_newton=4)
    reqs = [small_request(i, seed=13) for i in range(4)]
    srv = native.Server(max_active=4, prep_threads=2, pcg_chunk=8, **opts)
    ids = [srv.submit(sc, sc.steps, sc.dt) for sc in reqs]
    res = []
    for i in ids:
        try:
            res.append(srv.wait(i, timeout=120.0))
        except native.SorsError as e:
            res.append(e.code)
    srv.close()
    fallbacks = 0
    for sc, got in zip(reqs, res):
        r = native.Rollout.from_scene(sc, precision=native.MIXED, **opts)
        code, nfb = native.OK, 0
        try:
            for _ in range(sc.steps):
                nfb += r.step_resident(sc.dt).fp64_fallbacks
        except native.SorsError as e:
            code = e.code
        ref = r.positions().reshape(-1, 3)
        r.close()
        if code != native.OK:
            say(sc.name, "failed alone with", code, "served:", got if not isinstance(got, tuple) else "ok")
            assert got == code
            continue
        q, rs = got
        say(sc.name, "fallbacks served", rs["fp64_fallbacks"], "alone", nfb)
        assert rs["fp64_fallbacks"] == nfb
        assert np.array_equal(q, ref)
        fallbacks += nfb
    assert fallbacks > 0


def test_serving_long_rollouts_with_lagged_refresh_match_alone():
    """300-step served rollouts with lagged coarse refreshes (coarse_every = 4,
    factorised on each rollout's side stream while other rollouts' kernels
    run) equal their stand-alone runs bit for bit."""
    coarse = dict(coarse_nodes=64, coarse_every=4)
    reqs = [mesh.serving_request(i, seed=17, n_lo=800, n_hi=2500, steps_lo=300, steps_hi=300) for i in range(3)]
    srv = native.Server(max_active=3, prep_threads=2, pcg_chunk=8, **SOLVER, **coarse)
    ids = [srv.submit(sc, sc.steps, sc.dt) for sc in reqs]
    out = [srv.wait(i, timeout=300.0) for i in ids]
    srv.close()
    for sc, (q, rs) in zip(reqs, out):
        ref = alone(sc, **coarse)
        say(sc.name, sc.n_nodes, sc.steps, "long served vs alone", np.abs(q - ref).max(), rs["pcg_iters"])
        assert rs["status"] == native.OK and rs["steps"] == 300
        assert np.array_equal(q, ref)
