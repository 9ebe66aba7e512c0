#!/usr/bin/env python3
"""Implicit Newton steps per second on one MI355X.

Workload: the scaled beam -- 200x40x40 cells split into 1,920,000 Kuhn tet4
(337,881 nodes, 1,013,643 DOFs), compressible Neo-Hookean with E drawn
log-uniform in [5e5, 5e6] Pa per cell (SplitMix64 seed 0), nu = 0.45,
rho = 1100, gravity -z, x = 0 clamped, initial velocity 0.5 m/s * x/L in +z,
backward Euler dt = 1e-3, Newton + Armijo line search to ||g|| < 1e-8 N, PCG
(rtol 3e-3, 3e-2 for each step's first solve; a solve aimed at the residual
criterion reuses the last tangent) preconditioned by block-Jacobi plus a
rigid-body coarse correction on 2048-node aggregates (`--coarse-nodes`,
refreshed every `--coarse-every` tangents).  A step is one call of
Solver::step on device-resident state.

The timed run uses mixed precision (FP32 element tangents / CSR / SpMV / PCG
vectors, FP64 state, residual, energies, dot products and coarse solve).  The
same steps are then repeated in all-FP64 and the final states compared, and in
mixed precision with block-Jacobi alone; the CPU baseline is one full implicit
step of the same workload on the FP64 oracle (Newton + Armijo, block-Jacobi
PCG at the same tolerances) timed end to end on the host cores.  Side legs:
hex8 beam, contact poke, adjoint of a muscle-driven leg, serving of a seeded
request stream.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
       python bench.py --profile-kernel spmv|assemble|copy|update [--reps R]
"""
import argparse
import json
import os
import sys
import time
import traceback

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "implicit steps/sec @1M tet (≤1e-6 err vs FP64); served solves/sec + p50/p99"
HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_scene(args):
    from sors_amd import mesh

    if args.small:
        return mesh.beam(nx=40, ny=8, nz=8, element=args.element, e_range=(5e5, 5e6), seed=0, steps=20,
                         name="beam40x8x8")
    return mesh.cantilever_large(element=args.element, seed=0)


def run_gpu(sc, precision, steps, warmup, profile, **options):
    from sors_amd import native

    r = native.Rollout.from_scene(sc, precision=precision, profile_kernels=int(profile), **options)
    for _ in range(warmup):
        r.step_resident(sc.dt)
    agg = {"newton_iters": 0, "pcg_iters": 0, "backtracks": 0, "spmv_ms": 0.0, "spmv_launches": 0,
           "assemble_kernel_ms": 0.0, "assemble_launches": 0, "ms_assemble": 0.0, "ms_solve": 0.0,
           "ms_linesearch": 0.0, "fp64_fallbacks": 0}
    t0 = time.perf_counter()
    for _ in range(steps):
        st = r.step_resident(sc.dt)
        for k in agg:
            agg[k] += getattr(st, k)
    elapsed = time.perf_counter() - t0
    q = r.positions()
    return r, q, elapsed, agg


def solver_opts(args):
    """Solver options shared by the beam, hex and contact legs."""
    return dict(pcg_rtol=args.pcg_rtol, pcg_rtol_first=args.pcg_rtol_first, reuse_tangent=args.reuse_tangent,
                g_tol=args.g_tol, predict=args.predictor, coarse_nodes=args.coarse_nodes,
                coarse_every=args.coarse_every, pcg_check_every=args.check_every)


def cpu_baseline(sc, opts, cores):
    """One full implicit step of the workload on the FP64 CPU oracle, timed
    end to end on `cores` host threads: Newton + Armijo from the constant-
    velocity start (as the GPU's first step), block-Jacobi PCG at the GPU
    run's tolerances (pcg_rtol, g_tol), SPD-projected tangents.  Returns the
    baseline record and the oracle's positions after the step (the reference
    for the GPU's first step).  The context and the CSR pattern are set up
    before the timed region: they are one-time set-up, not part of a step."""
    from oracle import oracle as O

    o = O.System.from_scene(sc, method=O.NEWTON_LS, nthreads=cores)
    q = sc.X.reshape(-1) + sc.dt * sc.v0 * (1 - sc.bcmask)
    o.pattern(sc.dt)
    t0 = time.perf_counter()
    q1, st = o.step(q, sc.dt, linear_solver=O.PCG, project=1, pcg_rtol=opts["pcg_rtol"], g_tol=opts["g_tol"],
                    nthreads=cores)
    el = time.perf_counter() - t0
    sample = (f"one full implicit step of the same workload (first step from rest) on the FP64 oracle, {cores} "
              f"OpenMP threads, timed end to end: {st.newton_iters} Newton iterations, {st.pcg_iters} block-Jacobi "
              f"PCG iterations (rtol {opts['pcg_rtol']:g}), {st.backtracks} backtracks; "
              f"gradient {st.t_grad:.2f}s, Hessian {st.t_hess:.2f}s, solve {st.t_solve:.2f}s, "
              f"line search {st.t_energy:.2f}s")
    rec = {"value": 1.0 / el, "unit": "steps/s", "cores": cores, "kind": "port", "sample": sample,
           "seconds_per_step": el, "newton_iters": st.newton_iters, "pcg_iters": int(st.pcg_iters)}
    return rec, q1


def serve_leg(args, opts):
    """Serving: the seeded request stream (beam / cube tet meshes, 2k-200k
    nodes, E log-uniform [1e5, 1e7], nu uniform [0.3, 0.48], 10-50 steps of
    dt = 1e-3) arriving as a Poisson stream at `serve_rate` per second,
    served mixed-precision with continuous batching under the HBM budget.
    Requests are submitted at their arrival times; latency runs from submit
    to result.  The sequential plain-FP64 comparator runs every
    `serve_seq_every`-th request alone on the GPU; the service time of every
    other request is predicted from a least-squares fit of service time to
    node-steps over those, and the whole stream is queued first come, first
    served at the same arrival times."""
    from sors_amd import mesh, native

    count, rate = args.serve, args.serve_rate
    t0 = time.perf_counter()
    reqs = [mesh.serving_request(i, seed=args.serve_seed) for i in range(count)]
    arrivals = mesh.serving_arrivals(count, rate=rate, seed=args.serve_seed)
    sub = list(range(0, count, max(1, args.serve_seq_every)))
    subset = set(sub)
    log(f"serving: {count} requests generated in {time.perf_counter() - t0:.1f}s, "
        f"{sum(sc.n_nodes for sc in reqs)} nodes, {sum(sc.steps for sc in reqs)} steps in total")
    srv = native.Server(hbm_budget_bytes=args.serve_budget, max_active=args.serve_max_active,
                        prep_threads=args.serve_prep_threads, pcg_chunk=args.serve_chunk,
                        precision=native.MIXED, **opts)
    ids = []
    start = time.perf_counter()
    for t, sc in zip(arrivals, reqs):
        dt_wait = t - (time.perf_counter() - start)
        if dt_wait > 0:
            time.sleep(dt_wait)
        ids.append(srv.submit(sc, sc.steps, sc.dt))
    submit_lag = time.perf_counter() - start - float(arrivals[-1])
    stats, kept = [], {}
    for k, i in enumerate(ids):
        q, st = srv.wait(i)
        stats.append(st)
        if k in subset:
            kept[k] = q
        if (k + 1) % max(1, count // 16) == 0:
            log(f"  served {k + 1}/{count} after {time.perf_counter() - start:.1f}s")
    wall = time.perf_counter() - start
    sst = srv.stats()
    srv.close()
    lat = np.array([st["latency_ms"] for st in stats])
    steps_total = sum(sc.steps for sc in reqs)
    out = {
        "value": count / wall, "unit": "solves/s", "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)), "steps_per_s": steps_total / wall, "wall_s": wall,
        "requests": count, "arrival_rate": rate,
        "sample": f"requests 0..{count - 1} of the seeded stream (seed {args.serve_seed})",
        "nodes_total": int(sum(sc.n_nodes for sc in reqs)), "steps_total": int(steps_total),
        "last_submit_lag_s": submit_lag,
        "queue_ms_p50": float(np.percentile([st["queue_ms"] for st in stats], 50)),
        "queue_ms_p99": float(np.percentile([st["queue_ms"] for st in stats], 99)),
        "prep_ms_p50": float(np.percentile([st["prep_ms"] for st in stats], 50)),
        "newton_iters": int(sum(st["newton_iters"] for st in stats)),
        "pcg_iters": int(sum(st["pcg_iters"] for st in stats)),
        "fp64_fallbacks": int(sum(st["fp64_fallbacks"] for st in stats)),
        "server": {k: sst[k] for k in ("peak_active", "peak_resident_bytes", "ticks", "batched_launches")},
        "options": {"max_active": args.serve_max_active, "prep_threads": args.serve_prep_threads,
                    "pcg_chunk": args.serve_chunk, "hbm_budget_bytes": args.serve_budget},
    }
    log(f"served: {out['value']:.2f} solves/s, p50 {out['p50_ms']:.0f} ms, p99 {out['p99_ms']:.0f} ms, "
        f"wall {wall:.1f}s, {sst}")
    if args.no_serve_seq:
        return out
    # sequential plain FP64-GPU: the subset measured alone, one after another
    service, errs, work = [], [], []
    for n_done, k in enumerate(sub):
        sc = reqs[k]
        ts = time.perf_counter()
        r = native.Rollout.from_scene(sc, precision=native.FP64, **opts)
        for _ in range(sc.steps):
            r.step_resident(sc.dt)
        q64 = r.positions().reshape(-1, 3)
        r.close()
        service.append(time.perf_counter() - ts)
        work.append(float(sc.n_nodes) * sc.steps)
        errs.append(float(np.abs(kept[k] - q64).max() / sc.bbox_diag))
        if (n_done + 1) % max(1, len(sub) // 4) == 0:
            log(f"  sequential fp64 {n_done + 1}/{len(sub)}: {sum(service):.1f}s of service")
    A = np.stack([np.ones(len(work)), np.array(work)], axis=1)
    coef, *_ = np.linalg.lstsq(A, np.array(service), rcond=None)
    pred = A @ coef
    r2 = 1.0 - float(((np.array(service) - pred) ** 2).sum() / ((np.array(service) - np.mean(service)) ** 2).sum())
    est = [service[sub.index(k)] if k in subset else max(0.0, coef[0] + coef[1] * sc.n_nodes * sc.steps)
           for k, sc in enumerate(reqs)]
    finish, done = [], 0.0
    for t, sv in zip(arrivals, est):
        done = max(done, t) + sv
        finish.append(done)
    lat64 = 1e3 * (np.array(finish) - arrivals)
    out["sequential_fp64_gpu"] = {
        "value": count / finish[-1], "unit": "solves/s", "p50_ms": float(np.percentile(lat64, 50)),
        "p99_ms": float(np.percentile(lat64, 99)), "measured_requests": len(sub),
        "service_s_measured": float(sum(service)), "service_s_total_est": float(sum(est)),
        "fit": {"s_per_request": float(coef[0]), "s_per_node_step": float(coef[1]), "r2": r2},
        "how": f"every {args.serve_seq_every}th request alone on the GPU in plain FP64 (create, steps, read back), "
               "one after another; the others' service times from the least-squares fit of service time to "
               "node-steps over those; latency = FCFS queue at the same arrival times"}
    out["vs_sequential_fp64"] = out["value"] / out["sequential_fp64_gpu"]["value"]
    out["max_pos_err_over_bbox_vs_fp64"] = max(errs)
    log(f"sequential fp64: {out['sequential_fp64_gpu']}; max err/bbox {max(errs):.3e}")
    return out


def contact_leg(args, opts):
    """Contact poke: 40^3-cell Kuhn tet cube (384,000 tets), E = 2e5, nu = 0.4,
    resting on the ground plane (penalty 1e5 N/m, Coulomb mu = 0.5), a rigid
    disk (r = 0.01 m) moving -z at 0.02 m/s with a seeded 1 mm lateral
    jitter; dt = 5e-4, mixed precision.  The first `--contact-check` steps
    are repeated in plain FP64 for the error, and unless --no-cpu
    `--contact-oracle-steps` steps from the FP64 state at step
    `--contact-oracle-from` are compared with the FP64 CPU oracle."""
    from sors_amd import mesh, native

    p = mesh.contact_poke(n=40, steps=args.contact_steps, seed=0)
    sc = p.scene
    cfg = dict(planes=[((0.0, 0.0, 0.0), (0.0, 0.0, 1.0))], disks=[(p.centers[0], (0.0, 0.0, -1.0), p.radius)],
               surface=p.surface, stiffness=p.stiffness, friction=p.friction, eps_v=p.eps_v)

    def run(precision, steps, start=None, j0=0, state=False):
        r = native.Rollout.from_scene(sc, precision=precision, **opts)
        r.set_contact(**cfg)
        if start is not None:
            r.set_state(*start)
        agg = {"newton_iters": 0, "pcg_iters": 0, "backtracks": 0, "fp64_fallbacks": 0}
        t0 = time.perf_counter()
        for k in range(j0, j0 + steps):
            r.set_disk(0, p.centers[k + 1], (0.0, 0.0, -1.0), p.radius)
            st = r.step_resident(sc.dt)
            for key in agg:
                agg[key] += getattr(st, key)
        el = time.perf_counter() - t0
        q = r.positions().reshape(-1, 3)
        qv = r.state() if state else None
        r.close()
        return (q, el, agg, qv) if state else (q, el, agg)

    q, el, agg = run(native.MIXED, sc.steps)
    c = p.centers[sc.steps]
    top = sc.X[:, 2] > sc.X[:, 2].max() - 1e-9
    under = top & (np.linalg.norm(q[:, :2] - c[:2], axis=1) < p.radius)
    out = {"value": sc.steps / el, "unit": "steps/s", "steps": sc.steps, "ms_per_step": 1e3 * el / sc.steps,
           "workload": f"{sc.name}: {sc.n_elems} tet4, {sc.n_nodes} nodes, {len(p.surface)} surface nodes; "
                       f"ground penalty 1e5 N/m, mu 0.5, disk r 0.01 m at 0.02 m/s, dt 5e-4, mixed",
           "newton_per_step": agg["newton_iters"] / sc.steps, "pcg_per_step": agg["pcg_iters"] / sc.steps,
           "fp64_fallbacks": agg["fp64_fallbacks"],
           "max_ground_penetration_m": float(max(0.0, -q[:, 2].min())),
           "indentation_m": float(sc.X[:, 2].max() - q[under, 2].min()) if under.any() else 0.0}
    log(f"contact: {out}")
    if args.contact_check > 0:
        k = min(args.contact_check, sc.steps)
        qm, elm, _ = run(native.MIXED, k)
        q64, el64, agg64 = run(native.FP64, k)
        out["check_steps"] = k
        out["max_pos_err_over_bbox_mixed_vs_fp64"] = float(np.abs(qm - q64).max() / sc.bbox_diag)
        out["check_steps_per_s"] = {"mixed": k / elm, "fp64": k / el64}
    if not args.no_cpu and args.contact_oracle_steps > 0:
        # steps against the FP64 CPU oracle (block-Jacobi PCG at the same
        # tolerances, the same contact, friction and disk path), from the
        # FP64 state after `--contact-oracle-from` steps (the disk pressing)
        from oracle import oracle as O

        j0 = min(args.contact_oracle_from, sc.steps)
        k = min(args.contact_oracle_steps, sc.steps - j0)
        start = run(native.FP64, j0, state=True)[3] if j0 > 0 else (sc.X.reshape(-1).copy(), np.zeros(3 * sc.n_nodes))
        cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
        o = O.System.from_scene(sc, method=O.NEWTON_LS, nthreads=cores)
        o.set_contact(**cfg)
        o.set_initial_deformation(start[0])
        o.set_initial_velocity(start[1])
        qo = start[0].copy()
        t0 = time.perf_counter()
        for j in range(j0, j0 + k):
            o.set_disk(0, p.centers[j + 1], (0.0, 0.0, -1.0), p.radius)
            qo, _ = o.step(qo, sc.dt, linear_solver=O.PCG, project=1, pcg_rtol=opts["pcg_rtol"], g_tol=opts["g_tol"],
                           nthreads=cores)
        el_o = time.perf_counter() - t0
        qm = run(native.MIXED, k, start=start, j0=j0)[0]
        q64 = run(native.FP64, k, start=start, j0=j0)[0]
        moved = float(np.abs(qo - start[0]).max() / sc.bbox_diag)
        out["vs_cpu_oracle"] = {
            "from_step": j0, "steps": k,
            "max_pos_err_over_bbox_mixed": float(np.abs(qm.reshape(-1) - qo).max() / sc.bbox_diag),
            "max_pos_err_over_bbox_fp64": float(np.abs(q64.reshape(-1) - qo).max() / sc.bbox_diag),
            "moved_over_bbox": moved, "oracle_s_per_step": el_o / k, "cores": cores}
        log(f"contact vs CPU oracle: {out['vs_cpu_oracle']}")
    return out


def adjoint_leg(args):
    """Differentiable rollout: fibre-actuated leg prism 0.2 x 0.03 x 0.03 m
    (14 x 14 x 100 Kuhn cells: 117,600 tets, 22,725 nodes), top clamped, foot
    over a frictionless ground plane (penalty 1e5 N/m), 32 muscle groups,
    sigma_max 2e5 Pa, 100 backward Euler steps of dt = 2e-3; loss = squared
    distance of the foot centroid to a seeded target.  Gradient w.r.t. the
    117,600 Young's moduli and the 100 x 32 activations by the adjoint with
    sqrt(T) = 10 checkpoints, FP64, two-level PCG (`--adjoint-coarse` nodes
    per aggregate, refreshed from every tangent); checked against central
    differences (h = 1e-6) along `--adjoint-dirs` seeded directions."""
    from sors_amd import adjoint, mesh

    leg = mesh.leg_prism(steps=args.adjoint_steps, seed=0)
    ro = adjoint.LegRollout(leg, coarse_nodes=args.adjoint_coarse)
    E = leg.scene.youngs.copy()
    t0 = time.perf_counter()
    fwd = ro.forward(E, leg.act)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    L, gE, gA = ro.gradient(E, leg.act, fwd=fwd)
    t_bwd = time.perf_counter() - t0
    out = {"workload": f"{leg.scene.name}: {leg.scene.n_elems} tet4, {leg.scene.n_nodes} nodes, "
                       f"{args.adjoint_steps} steps dt 2e-3, 32 muscle groups, FP64, minimize_newton",
           "loss": L, "forward_s": t_fwd, "gradient_s": t_bwd,
           "gradient_over_forward": t_bwd / t_fwd, "params": int(gE.size + gA.size),
           "forward_newton": ro.stats["forward_newton"], "forward_pcg": ro.stats["forward_pcg"],
           "adjoint_pcg": ro.stats["adjoint_pcg"], "checkpoint_every": ro.K,
           "coarse_nodes": args.adjoint_coarse}
    log(f"adjoint: {out}")
    if args.adjoint_dirs > 0:
        rng = np.random.default_rng(7)
        h = 1e-6
        jobs, ads = [], []
        for k in range(args.adjoint_dirs):
            dE = rng.normal(size=E.shape)
            dA = rng.normal(size=leg.act.shape)
            nrm = np.sqrt((dE * dE).sum() + (dA * dA).sum())
            dE /= nrm
            dA /= nrm
            jobs += [(E * (1 + h * dE), leg.act + h * dA), (E * (1 - h * dE), leg.act - h * dA)]
            ads.append(float((gE * E) @ dE) + float((gA * dA).sum()))
        t0 = time.perf_counter()
        losses = adjoint.perturbed_losses(ro, jobs, workers=args.adjoint_workers)
        t_fd = time.perf_counter() - t0
        checks = []
        for k, ad in enumerate(ads):
            fd = (losses[2 * k] - losses[2 * k + 1]) / (2 * h)
            checks.append({"fd": fd, "adjoint": ad, "rel_err": abs(fd - ad) / max(abs(fd), abs(ad), 1e-300)})
            log(f"  direction {k}: fd {fd:.9e} adjoint {ad:.9e}")
        out["fd_check"] = {"h": h, "directions": args.adjoint_dirs, "max_rel_err": max(c["rel_err"] for c in checks),
                           "checks": checks, "seconds": t_fd,
                           "how": f"{2 * args.adjoint_dirs} perturbed {args.adjoint_steps}-step rollouts (two per direction), "
                                  f"{args.adjoint_workers} at a time on their own device contexts and streams"}
    ro.close()
    return out


def hex_leg(args, opts):
    """The same beam physics on the 200 x 40 x 40 hex8 grid (320,000
    elements, 2 x 2 x 2 Gauss points, 27-point node stencil), mixed
    precision, against plain FP64 on the GPU for the same steps; unless
    --no-cpu, both precisions' first step from rest against one full step of
    the FP64 CPU oracle."""
    from sors_amd import mesh, native

    sc = mesh.cantilever_large(element="hex", seed=0)
    out = {"workload": f"{sc.name}: {sc.n_elems} hex8, {sc.n_nodes} nodes, E log-uniform [5e5,5e6] seed 0, "
                       f"nu 0.45, BE dt 1e-3", "steps": args.hex_steps}
    res = {}
    for name, prec in (("mixed", native.MIXED), ("fp64", native.FP64)):
        r, q, el, agg = run_gpu(sc, prec, args.hex_steps, 1, profile=False, **opts)
        r.close()
        res[name] = q
        out[name] = {"steps_per_s": args.hex_steps / el, "ms_per_step": 1e3 * el / args.hex_steps,
                     "newton_per_step": agg["newton_iters"] / args.hex_steps,
                     "pcg_per_step": agg["pcg_iters"] / args.hex_steps}
        log(f"hex {name}: {out[name]}")
    out["max_pos_err_over_bbox_mixed_vs_fp64"] = float(np.abs(res["mixed"] - res["fp64"]).max() / sc.bbox_diag)
    if not args.no_cpu:
        # the first step from rest against one full step of the FP64 CPU oracle
        cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
        cpu, q_cpu = cpu_baseline(sc, opts, cores)
        first = {}
        for name, prec in (("mixed", native.MIXED), ("fp64", native.FP64)):
            r1, q1, _, _ = run_gpu(sc, prec, 1, 0, profile=False, **opts)
            r1.close()
            first[name] = float(np.abs(q1 - q_cpu).max() / sc.bbox_diag)
        out["vs_cpu_oracle_first_step"] = {
            "max_pos_err_over_bbox_mixed": first["mixed"], "max_pos_err_over_bbox_fp64": first["fp64"],
            "moved_over_bbox": float(np.abs(q_cpu - sc.X.reshape(-1)).max() / sc.bbox_diag),
            "oracle_seconds": cpu["seconds_per_step"], "oracle_newton": cpu["newton_iters"],
            "oracle_pcg": cpu["pcg_iters"], "cores": cores}
        log(f"hex first step vs CPU oracle: {out['vs_cpu_oracle_first_step']}")
    return out


def load_traffic(key):
    try:
        with open(TRAFFIC_FILE) as f:
            return json.load(f).get(key)
    except (OSError, ValueError):
        return None


def profile_kernel(args):
    """Launch one kernel `reps` times on an assembled system (for rocprofv3 runs)."""
    from sors_amd import native

    sc = make_scene(args)
    prec = native.FP64 if args.fp64 else native.MIXED
    r = native.Rollout.from_scene(sc, precision=prec)
    rng = np.random.default_rng(1)
    q = sc.X.reshape(-1) + 1e-5 * rng.normal(size=3 * sc.n_nodes) * (1 - sc.bcmask)
    r.set_initial_deformation(q)
    r.linear_solve(q, sc.dt, np.ones(3 * sc.n_nodes), rtol=0.0, max_iter=2)
    kind = {"spmv": native.KERNEL_SPMV, "assemble": native.KERNEL_ASSEMBLE, "copy": native.KERNEL_COPY,
            "update": native.KERNEL_PCG_UPDATE}[args.profile_kernel]
    ms, by = r.bench_kernel(kind, args.reps)
    print(json.dumps({"kernel": args.profile_kernel, "avg_ms": ms, "algo_bytes": by,
                      "GBps": by / (ms * 1e-3) / 1e9, "reps": args.reps}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--element", choices=["tet", "hex"], default="tet")
    ap.add_argument("--small", action="store_true", help="40x8x8 cells instead of 200x40x40 (debug)")
    ap.add_argument("--no-fp64", action="store_true", help="skip the all-FP64 comparison run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline sample")
    ap.add_argument("--profile-kernel", choices=["spmv", "assemble", "copy", "update"])
    ap.add_argument("--reps", type=int, default=50)
    ap.add_argument("--fp64", action="store_true", help="profile-kernel in FP64")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch PCG kernels directly instead of replaying captured hipGraphs (profiler runs)")
    ap.add_argument("--pcg-rtol", type=float, default=3e-3)
    ap.add_argument("--pcg-rtol-first", type=float, default=3e-2,
                    help="PCG tolerance of each step's first Newton solve (0: --pcg-rtol)")
    ap.add_argument("--reuse-tangent", type=int, default=1, choices=[0, 1],
                    help="a solve aimed at g_tol may reuse the last tangent (options.reuse_tangent)")
    ap.add_argument("--g-tol", type=float, default=1e-8, help="Newton residual criterion ||g||_2 (N)")
    ap.add_argument("--no-predict", action="store_true", help="start Newton from the previous positions")
    ap.add_argument("--predictor", type=int, default=2, choices=[0, 1, 2],
                    help="Newton start: 0 previous positions, 1 constant velocity, 2 constant acceleration")
    ap.add_argument("--check-every", type=int, default=16,
                    help="PCG iterations before a solve's first host convergence check")
    ap.add_argument("--coarse-nodes", type=int, default=2048,
                    help="nodes per aggregate of the two-level PCG preconditioner (0: block-Jacobi only)")
    ap.add_argument("--coarse-every", type=int, default=16, help="coarse operator refreshed every k-th tangent")
    ap.add_argument("--serve", type=int, default=2048, help="serving requests to run (0: skip the serving leg)")
    ap.add_argument("--serve-only", action="store_true", help="run only the serving leg")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-seed", type=int, default=0)
    ap.add_argument("--serve-budget", type=float, default=200e9, help="HBM budget of resident rollouts (bytes)")
    ap.add_argument("--serve-max-active", type=int, default=64)
    ap.add_argument("--serve-prep-threads", type=int, default=8)
    ap.add_argument("--serve-chunk", type=int, default=8)
    ap.add_argument("--no-serve-seq", action="store_true", help="skip the sequential FP64-GPU serving baseline")
    ap.add_argument("--serve-seq-every", type=int, default=32,
                    help="the sequential FP64 comparator measures every n-th request")
    ap.add_argument("--contact-steps", type=int, default=500, help="contact poke steps (0: skip the contact leg)")
    ap.add_argument("--contact-check", type=int, default=500, help="contact steps repeated in FP64 for the error")
    ap.add_argument("--contact-only", action="store_true", help="run only the contact leg")
    ap.add_argument("--contact-oracle-steps", type=int, default=5,
                    help="contact steps compared with the FP64 CPU oracle (0: skip)")
    ap.add_argument("--contact-oracle-from", type=int, default=100,
                    help="step (of the FP64 GPU rollout) the oracle comparison starts from")
    ap.add_argument("--adjoint-steps", type=int, default=100, help="leg rollout steps (0: skip the adjoint leg)")
    ap.add_argument("--adjoint-dirs", type=int, default=20, help="finite-difference directions checked")
    ap.add_argument("--adjoint-workers", type=int, default=8,
                    help="perturbed rollouts of the finite-difference check run at once")
    ap.add_argument("--adjoint-only", action="store_true", help="run only the adjoint leg")
    ap.add_argument("--adjoint-coarse", type=int, default=256,
                    help="nodes per aggregate of the adjoint leg's two-level PCG (0: block-Jacobi)")
    ap.add_argument("--hex-steps", type=int, default=20, help="hex8 beam steps (0: skip the hex leg)")
    ap.add_argument("--hex-only", action="store_true", help="run only the hex8 leg")
    args = ap.parse_args()
    if args.no_predict:
        args.predictor = 0
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)
    if args.profile_kernel:
        profile_kernel(args)
        return
    from sors_amd import native

    if args.hex_only:
        print(json.dumps({"metric": METRIC, "hex": hex_leg(args, solver_opts(args))}), flush=True)
        return
    if args.adjoint_only:
        print(json.dumps({"metric": METRIC, "adjoint": adjoint_leg(args)}), flush=True)
        return
    if args.contact_only:
        print(json.dumps({"metric": METRIC, "contact": contact_leg(args, solver_opts(args))}), flush=True)
        return
    if args.serve_only:
        print(json.dumps({"metric": METRIC, "serving": serve_leg(args, solver_opts(args))}), flush=True)
        return
    sc = make_scene(args)
    log(f"workload {sc.name}: {sc.n_elems} elements, {sc.n_nodes} nodes; warmup {args.warmup}, steps {args.steps}")
    opts = dict(solver_opts(args), use_graph=int(not args.no_graph), pcg_check_every=args.check_every)
    r, q_mixed, elapsed, agg = run_gpu(sc, native.MIXED, args.steps, args.warmup, profile=True, **opts)
    nnz = int(len(r.pattern(sc.dt)[1]))
    ms_per_step = 1e3 * elapsed / args.steps
    log(f"mixed: {ms_per_step:.1f} ms/step, {agg}")
    spmv_ms = agg["spmv_ms"] / max(1, agg["spmv_launches"])
    n = 3 * sc.n_nodes
    # compulsory bytes of one SpMV launch for the layout in use (values once,
    # one column index (or 16-bit offset) per nonzero 3x3 node block, x and y once)
    iso_ms, spmv_bytes = r.bench_kernel(native.KERNEL_SPMV, reps=20)
    achieved = spmv_bytes / (spmv_ms * 1e-3) / 1e9 if spmv_ms > 0 else 0.0
    r.close()
    out = {
        "metric": METRIC,
        "value": args.steps / elapsed,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"{sc.name}: {sc.n_elems} {'tet4' if sc.ed == 4 else 'hex8'}, {sc.n_nodes} nodes, "
                        f"{n} DOFs, Hessian nnz {nnz}; E log-uniform [5e5,5e6] seed 0, nu 0.45, BE dt 1e-3, "
                        f"Newton + Armijo (||g|| < {args.g_tol:g} N or ||dx|| < 1e-9 m), PCG rtol {args.pcg_rtol:g} "
                        f"({args.pcg_rtol_first:g} for a step's first solve{', tangent reuse' if args.reuse_tangent else ''}) "
                        f"preconditioned by block-Jacobi + rigid-body coarse correction ({args.coarse_nodes} nodes "
                        f"per aggregate, refreshed every {args.coarse_every} tangents), mixed FP32/FP64",
            "elements": sc.n_elems, "nodes": sc.n_nodes, "nnz": nnz, "precision": "mixed",
        },
        "roofline": {
            "bound": "hbm", "kernel": "k_pcg_spmv_blk (sliced 3x3 node-block SpMV + w.u)", "achieved": achieved,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": load_traffic("spmv_tet_mixed"), "algo_bytes_per_launch": spmv_bytes,
            "avg_launch_ms": spmv_ms, "launches": agg["spmv_launches"], "isolated_avg_launch_ms": iso_ms,
        },
        "solver": {k: (v / args.steps) for k, v in agg.items()},
    }
    if agg["assemble_launches"]:
        asm_ms = agg["assemble_kernel_ms"] / agg["assemble_launches"]
        asm_bytes = 4.0 * sc.ed * sc.n_elems + 16.0 * sc.n_elems + 48.0 * sc.n_nodes + 4.0 * nnz + 8.0 * n
        out["assembly"] = {"avg_launch_ms": asm_ms, "algo_bytes_per_launch": asm_bytes,
                           "achieved_GBps": asm_bytes / (asm_ms * 1e-3) / 1e9}
    if not args.no_fp64:
        # plain FP64-GPU with the same solver settings, and a tightly converged
        # FP64 reference (PCG rtol 1e-9, ||g|| < 1e-11, ||dx|| < 1e-12) for the error
        r64, q64, el64, agg64 = run_gpu(sc, native.FP64, args.steps, args.warmup, profile=False, **opts)
        r64.close()
        tight = dict(opts, pcg_rtol=1e-9, g_tol=1e-11, dx_tol=1e-12)
        rref, qref, elref, _ = run_gpu(sc, native.FP64, args.steps, args.warmup, profile=False, **tight)
        rref.close()
        err_mixed = float(np.abs(q_mixed - qref).max() / sc.bbox_diag)
        err64 = float(np.abs(q64 - qref).max() / sc.bbox_diag)
        out["error"] = {"max_pos_err_over_bbox_mixed": err_mixed, "max_pos_err_over_bbox_fp64": err64,
                        "reference": "FP64 GPU path converged to PCG rtol 1e-9, ||g|| < 1e-11 N, ||dx|| < 1e-12 m "
                                     "(parity-tested against the CPU FP64 oracle on the 20x4x4 beam)",
                        "bound": 1e-6}
        out["fp64_gpu"] = {"value": args.steps / el64, "unit": "steps/s", "ms_per_step": 1e3 * el64 / args.steps,
                           "newton_per_step": agg64["newton_iters"] / args.steps,
                           "pcg_per_step": agg64["pcg_iters"] / args.steps}
        log(f"fp64: {1e3 * el64 / args.steps:.1f} ms/step; error/bbox mixed {err_mixed:.3e} fp64 {err64:.3e}")
    if not args.no_cpu:
        # like for like with the oracle's PCG (block-Jacobi alone): the same
        # steps on the GPU with coarse_nodes=0
        rbj, _, elbj, aggbj = run_gpu(sc, native.MIXED, args.steps, args.warmup, profile=False,
                                      **dict(opts, coarse_nodes=0))
        rbj.close()
        out["block_jacobi_gpu"] = {"value": args.steps / elbj, "unit": "steps/s",
                                   "ms_per_step": 1e3 * elbj / args.steps,
                                   "newton_per_step": aggbj["newton_iters"] / args.steps,
                                   "pcg_per_step": aggbj["pcg_iters"] / args.steps}
        log(f"block-Jacobi only: {out['block_jacobi_gpu']}")
        cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
        cpu, q_cpu = cpu_baseline(sc, solver_opts(args), cores)
        out["cpu_baseline"] = cpu
        out["cpu_baseline"]["gpu_over_cpu"] = out["value"] / cpu["value"]
        out["cpu_baseline"]["block_jacobi_gpu_over_cpu"] = out["block_jacobi_gpu"]["value"] / cpu["value"]
        log(f"cpu baseline: {out['cpu_baseline']}")
        # the GPU's first step from rest, mixed and plain FP64, against the oracle's
        first = {}
        for name, prec in (("mixed", native.MIXED), ("fp64", native.FP64)):
            r1, q1, _, _ = run_gpu(sc, prec, 1, 0, profile=False, **opts)
            r1.close()
            first[name] = float(np.abs(q1 - q_cpu).max() / sc.bbox_diag)
        out.setdefault("error", {})["vs_cpu_oracle_first_step"] = {
            "max_pos_err_over_bbox_mixed": first["mixed"], "max_pos_err_over_bbox_fp64": first["fp64"],
            "reference": "FP64 CPU oracle, one full step of the same workload, same start and tolerances",
            "moved_over_bbox": float(np.abs(q_cpu - sc.X.reshape(-1)).max() / sc.bbox_diag)}
        log(f"first step vs CPU oracle: {out['error']['vs_cpu_oracle_first_step']}")
    def leg(name, run):
        # a leg that fails (a solver error such as a Newton step that does not
        # converge, as the reference raises it) is reported in its field; the
        # headline line is still printed
        try:
            out[name] = run()
        except Exception as exc:  # noqa: BLE001 - any leg failure is recorded, not fatal
            traceback.print_exc()
            out[name] = {"error": f"{type(exc).__name__}: {exc}"}
            log(f"{name} leg failed: {exc}")

    if args.hex_steps > 0:
        leg("hex", lambda: hex_leg(args, solver_opts(args)))
    if args.contact_steps > 0:
        leg("contact", lambda: contact_leg(args, solver_opts(args)))
    if args.adjoint_steps > 0:
        leg("adjoint", lambda: adjoint_leg(args))
    if args.serve > 0:
        leg("serving", lambda: serve_leg(args, solver_opts(args)))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
