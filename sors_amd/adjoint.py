"""Differentiable implicit rollouts: the adjoint of a sequence of backward
Euler steps, with sqrt(T) checkpointing, behind a torch.autograd.Function.

The reference has no gradients (its optimisation is gradient-free); this is
the device path's addition.  Forward: T implicit steps on the device with
per-step activations act[t] (sors_set_activation) and per-element Young's
moduli E (sors_set_youngs); the loss is |foot(q_T) - target|^2 with foot the
centroid of a node set.  Backward: states are kept only every K ~ sqrt(T)
steps; each segment is recomputed from its checkpoint, then walked backwards
with sors_adjoint_step, which solves one transposed system per step and
accumulates dL/dE and dL/dact on the device.

Every solve, forward and backward, runs through the HIP library
(sors_amd/libsors_amd.so); there is no host fallback.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import native


class LegRollout:
    """One device rollout of a LegScene (mesh.leg_prism) with its
    actuation, contact and loss."""

    def __init__(self, leg, precision=native.FP64, checkpoint_every=None, adjoint_rtol=1e-11,
                 method=native.MINIMIZE_NEWTON, coarse_nodes=0, **solver):
        """Solves use minimize_newton (solver.cpp:56-105) on the unprojected
        Hessian by default: its steps converge quadratically to rounding
        (about 5 per step here), whereas an Armijo test stops resolving energy
        decreases near 1e-16 of the energy and the SPD-projected tangent makes
        Newton converge only linearly; either would leave forward states (and
        finite differences of the loss) noisy at the solver tolerance."""
        self.leg = leg
        self._init = dict(precision=precision, checkpoint_every=checkpoint_every, adjoint_rtol=adjoint_rtol,
                          method=method, coarse_nodes=coarse_nodes, **solver)
        sc = leg.scene
        self.sc = sc
        self.T = sc.steps
        self.K = checkpoint_every or max(1, int(round(math.sqrt(self.T))))
        self.adjoint_rtol = adjoint_rtol
        # predict=1: Newton's start depends on the state only (not on the step
        # before it), so recomputing from any checkpoint reproduces the forward
        # states bit for bit; coarse_every=1 for the same reason (a two-level
        # preconditioner, when asked for, refreshed from every tangent)
        opts = dict(g_tol=0.0, dx_tol=1e-12, pcg_rtol=1e-12, max_newton=100, project_spd=0, predict=1,
                    coarse_nodes=coarse_nodes, coarse_every=1)
        opts.update(solver)
        self.r = native.Rollout.from_scene(sc, method=method, precision=precision, **opts)
        self.r.set_contact(planes=[((0.0, 0.0, 0.0), (0.0, 0.0, 1.0))], surface=leg.surface,
                           stiffness=leg.stiffness, friction=0.0)
        self.r.set_active(leg.group, leg.fibre, leg.sigma_max)
        self.n_groups = int(leg.group.max()) + 1
        self.q0 = sc.X.reshape(-1).copy()
        self.v0 = np.zeros(3 * sc.n_nodes)
        self.stats = {}

    def close(self):
        self.r.close()

    def clone(self):
        """A second rollout of the same leg and options on a context (and
        stream) of its own: its forward runs are bit-identical to this one's."""
        return LegRollout(self.leg, **self._init)

    # ---- loss -------------------------------------------------------------
    def loss_of(self, q):
        foot = q.reshape(-1, 3)[self.leg.foot].mean(axis=0)
        d = foot - self.leg.target
        return float(d @ d)

    def loss_grad_q(self, q):
        foot = q.reshape(-1, 3)[self.leg.foot].mean(axis=0)
        g = np.zeros((self.sc.n_nodes, 3))
        g[self.leg.foot] = 2.0 * (foot - self.leg.target) / len(self.leg.foot)
        return g.reshape(-1)

    # ---- forward ----------------------------------------------------------
    def _run(self, E, act, t0, t1, q, v, keep=None):
        """Steps t0..t1-1 from state (q, v); keep: dict t -> (q_t, v_t) for t in
        the range (after each step) when given."""
        r = self.r
        r.set_youngs(E)
        r.set_state(q, v)
        newton = pcg = 0
        for t in range(t0, t1):
            r.set_activation(act[t])
            st = r.step_resident(self.sc.dt)
            newton += st.newton_iters
            pcg += st.pcg_iters
            if keep is not None:
                keep[t + 1] = r.state()
        qT, vT = r.state()
        return qT, vT, newton, pcg

    def forward(self, E, act):
        """Loss and the checkpoints {t: (q_t, v_t)} for t = 0, K, 2K, ..."""
        E = np.ascontiguousarray(E, dtype=np.float64)
        act = np.ascontiguousarray(act, dtype=np.float64).reshape(self.T, self.n_groups)
        ck = {0: (self.q0.copy(), self.v0.copy())}
        q, v = self.q0, self.v0
        t0 = time.perf_counter()
        newton = pcg = 0
        for s in range(0, self.T, self.K):
            e = min(self.T, s + self.K)
            q, v, nw, pc = self._run(E, act, s, e, q, v)
            newton += nw
            pcg += pc
            if e < self.T:
                ck[e] = (q.copy(), v.copy())
        self.stats["forward_s"] = time.perf_counter() - t0
        self.stats["forward_newton"] = newton
        self.stats["forward_pcg"] = pcg
        return self.loss_of(q), q, ck

    def loss(self, E, act):
        return self.forward(E, act)[0]

    # ---- backward ---------------------------------------------------------
    def gradient(self, E, act, fwd=None):
        """(loss, dL/dE [E], dL/dact [T, n_groups])."""
        E = np.ascontiguousarray(E, dtype=np.float64)
        act = np.ascontiguousarray(act, dtype=np.float64).reshape(self.T, self.n_groups)
        L, qT, ck = fwd if fwd is not None else self.forward(E, act)
        t0 = time.perf_counter()
        Ebar = np.zeros(self.sc.n_elems)
        abar = np.zeros((self.T, self.n_groups))
        qbar = self.loss_grad_q(qT)
        vbar = np.zeros_like(qbar)
        dt = self.sc.dt
        adj_pcg = 0
        starts = sorted(ck)
        for s in reversed(starts):
            e = min(self.T, s + self.K)
            keep = {s: ck[s]}
            self._run(E, act, s, e, ck[s][0], ck[s][1], keep=keep)  # recompute the segment
            for t in range(e - 1, s - 1, -1):
                q_prev, v_prev = keep[t]
                q_next = keep[t + 1][0]
                self.r.set_activation(act[t])
                row = np.zeros(self.n_groups)
                qbar, vbar, it = self.r.adjoint_step(q_prev, v_prev, q_next, dt, qbar, vbar, youngs_bar=Ebar,
                                                     act_bar=row, rtol=self.adjoint_rtol)
                abar[t] += row
                adj_pcg += it
        self.stats["backward_s"] = time.perf_counter() - t0
        self.stats["adjoint_pcg"] = adj_pcg
        return L, Ebar, abar


def _torch_function():
    import torch

    class RolloutLoss(torch.autograd.Function):
        """loss = RolloutLoss.apply(E, act, rollout): the foot-to-target loss of a
        LegRollout as a differentiable function of per-element Young's moduli E
        [n_elems] and activations act [T, n_groups] (float64 tensors)."""

        @staticmethod
        def forward(ctx, E, act, rollout):
            En = E.detach().cpu().numpy().astype(np.float64)
            An = act.detach().cpu().numpy().astype(np.float64)
            fwd = rollout.forward(En, An)
            ctx.rollout = rollout
            ctx.saved = (En, An, fwd)
            ctx.dev = (E.device, act.device)
            return torch.tensor(fwd[0], dtype=torch.float64, device=E.device)

        @staticmethod
        def backward(ctx, gout):
            En, An, fwd = ctx.saved
            _, Ebar, abar = ctx.rollout.gradient(En, An, fwd=fwd)
            g = float(gout)
            return (torch.from_numpy(Ebar * g).to(ctx.dev[0]), torch.from_numpy(abar * g).to(ctx.dev[1]), None)

    return RolloutLoss


_FN = None


def rollout_loss(E, act, rollout):
    """Differentiable loss of a LegRollout (torch tensors in, scalar tensor out)."""
    global _FN
    if _FN is None:
        _FN = _torch_function()
    return _FN.apply(E, act, rollout)


def fd_check(rollout, E, act, directions=3, h=1e-6, seed=0, workers=1):
    """Central differences of the loss along seeded directions in (E, act)
    space (E perturbed relatively: E * (1 + h d_E)), against the adjoint
    gradient.  Returns a list of (fd, adjoint) directional derivatives.

    workers > 1: the 2 x directions perturbed rollouts run on that many
    clones of `rollout` (each its own device context and stream) from as many
    host threads, so their kernels overlap on the device; every rollout is
    deterministic, so the losses are those of a sequential run."""
    rng = np.random.default_rng(seed)
    L, gE, gA = rollout.gradient(E, act)
    jobs, ads = [], []
    for _ in range(directions):
        dE = rng.normal(size=E.shape)
        dA = rng.normal(size=act.shape)
        nrm = math.sqrt(float(dE @ dE) + float((dA * dA).sum()))
        dE /= nrm
        dA /= nrm
        jobs += [(E * (1 + h * dE), act + h * dA), (E * (1 - h * dE), act - h * dA)]
        ads.append(float((gE * E) @ dE) + float((gA * dA).sum()))
    losses = perturbed_losses(rollout, jobs, workers)
    out = [((losses[2 * k] - losses[2 * k + 1]) / (2 * h), ads[k]) for k in range(directions)]
    return L, out


def perturbed_losses(rollout, jobs, workers=1):
    """Losses of the rollouts (E, act) in `jobs`, in order; workers > 1 runs
    them on that many clones of `rollout` from as many host threads."""
    if workers <= 1 or len(jobs) <= 1:
        return [rollout.loss(Ej, Aj) for Ej, Aj in jobs]
    import queue
    from concurrent.futures import ThreadPoolExecutor

    n = min(workers, len(jobs))
    pool = queue.Queue()
    clones = [rollout.clone() for _ in range(n)]
    for c in clones:
        pool.put(c)

    def run(job):
        c = pool.get()
        try:
            return c.loss(*job)
        finally:
            pool.put(c)

    try:
        with ThreadPoolExecutor(n) as ex:
            return list(ex.map(run, jobs))
    finally:
        for c in clones:
            c.close()
