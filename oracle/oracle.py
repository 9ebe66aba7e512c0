"""ctypes binding of the CPU FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (sors_amd) never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

BE, CN = 0, 1
NEWTON, NEWTON_LS, SQP, FROZEN = 0, 1, 2, 3
LDLT, PCG = 0, 1
MAT_NH, MAT_SNH = 0, 1
TEST_POLY = -4
ERRORS = {
    -1: "Inverted Elements",
    -2: "Newton's method did not converge",
    -3: "Newton's method diverged",
    -4: "Line search failed",
    -5: "bad argument",
    -6: "linear solve failed",
}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Opts(C.Structure):
    _fields_ = [
        ("linear_solver", C.c_int),
        ("project", C.c_int),
        ("max_iter", C.c_int),
        ("dx_tol", C.c_double),
        ("g_tol", C.c_double),
        ("pcg_rtol", C.c_double),
        ("pcg_max_iter", C.c_int),
        ("max_backtrack", C.c_int),
        ("nthreads", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("newton_iters", C.c_int),
        ("pcg_iters", C.c_longlong),
        ("backtracks", C.c_int),
        ("gnorm", C.c_double),
        ("dxnorm", C.c_double),
        ("t_grad", C.c_double),
        ("t_hess", C.c_double),
        ("t_solve", C.c_double),
        ("t_energy", C.c_double),
    ]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _ip, _dp, _dp, _dp, _ip, _dp, _dp, C.c_double,
                                 C.c_int, C.c_int]
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_energy.restype = C.c_double
        L.orc_energy.argtypes = [C.c_void_p, _dp, C.c_double]
        L.orc_gradient.argtypes = [C.c_void_p, _dp, C.c_double, _dp]
        L.orc_pattern.restype = C.c_longlong
        L.orc_pattern.argtypes = [C.c_void_p, C.c_double]
        L.orc_get_pattern.argtypes = [C.c_void_p, _ip, _ip]
        L.orc_get_scatter_map.argtypes = [C.c_void_p, _ip]
        L.orc_hessian.argtypes = [C.c_void_p, _dp, C.c_double, C.c_int, _dp]
        L.orc_step.argtypes = [C.c_void_p, _dp, C.c_double, C.c_int, C.POINTER(Opts), C.POINTER(Stats)]
        L.orc_minimize.argtypes = [C.c_void_p, _dp, C.c_double, C.POINTER(Opts), C.POINTER(Stats)]
        L.orc_update_state.argtypes = [C.c_void_p, _dp, C.c_double]
        L.orc_get_mass.argtypes = [C.c_void_p, _dp]
        L.orc_set_initial_velocity.argtypes = [C.c_void_p, _dp]
        L.orc_set_initial_deformation.argtypes = [C.c_void_p, _dp]
        L.orc_get_state.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_set_threads.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_external_force.argtypes = [C.c_void_p, _dp]
        L.orc_default_opts.argtypes = [C.POINTER(Opts)]
        L.orc_elem_eval.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_double, C.c_int, _dp, _dp, _dp]
        L.orc_tet_precompute.argtypes = [_dp, _dp, _dp, _dp]
        L.orc_hex_precompute.argtypes = [_dp, _dp, _dp, _dp]
        L.orc_nh_psi.restype = C.c_double
        L.orc_nh_psi.argtypes = [_dp, C.c_double, C.c_double]
        L.orc_nh_dpsi.argtypes = [_dp, C.c_double, C.c_double, _dp]
        L.orc_nh_d2psi.argtypes = [_dp, C.c_double, C.c_double, _dp]
        L.orc_project_psd.argtypes = [C.c_int, _dp]
        L.orc_bench_pcg.restype = C.c_double
        L.orc_bench_pcg.argtypes = [C.c_void_p, _dp, C.c_double, C.c_int, C.c_int]
        L.orc_set_contact.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_int, _dp, _dp, _dp, C.c_int, _ip,
                                      C.c_double, C.c_double, C.c_double]
        L.orc_set_disk.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_double]
        L.orc_contact_energy.argtypes = [C.c_void_p, _dp, C.c_double]
        L.orc_contact_energy.restype = C.c_double
        L.orc_contact_lag.argtypes = [C.c_void_p, C.c_double]
        L.orc_contact_lag.restype = None
        L.orc_set_active.argtypes = [C.c_void_p, C.c_int, _ip, _dp, C.c_double]
        L.orc_set_activation.argtypes = [C.c_void_p, _dp]
        L.orc_set_activation.restype = None
        L.orc_set_youngs.argtypes = [C.c_void_p, _dp]
        L.orc_set_youngs.restype = None
        L.orc_snh_psi.restype = C.c_double
        L.orc_snh_psi.argtypes = [_dp, C.c_double, C.c_double]
        L.orc_snh_dpsi.argtypes = [_dp, C.c_double, C.c_double, _dp]
        L.orc_snh_dpsi.restype = None
        L.orc_snh_d2psi.argtypes = [_dp, C.c_double, C.c_double, _dp]
        L.orc_snh_d2psi.restype = None
        L.orc_elem_eval_ex.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_double, C.c_int, _dp, C.c_double,
                                       C.c_double, C.c_int, _dp, _dp, _dp]
        L.orc_set_material.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_muscle.argtypes = [C.c_void_p, C.c_int, _ip, _dp, _dp]
        L.orc_set_muscle_activation.argtypes = [C.c_void_p, _dp]
        L.orc_set_muscle_activation.restype = None
        L.orc_set_pressure.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, _ip, C.c_int]
        L.orc_set_pressure_activation.argtypes = [C.c_void_p, _dp]
        L.orc_set_pressure_activation.restype = None
        L.orc_pressure_force.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_pressure_force.restype = None
        L.orc_pressure_jacobian.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_pressure_jacobian.restype = None
        L.orc_qp_solve.argtypes = [C.c_int, _ip, _ip, _dp, _dp, C.c_int, _ip, _dp, _dp, _dp, _dp, _dp,
                                   C.POINTER(C.c_int)]
        L.orc_sqp_active.argtypes = [C.c_void_p, _dp, _ip, _ip, C.POINTER(C.c_int)]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _i(a):
    return a.ctypes.data_as(_ip) if a is not None else None


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def opts(**kw) -> Opts:
    o = Opts()
    lib().orc_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


# ---- element level -------------------------------------------------------
def tet_precompute(Xe):
    Xe = f64(Xe).reshape(12)
    Dminv, V0, B = np.zeros(9), np.zeros(1), np.zeros(108)
    lib().orc_tet_precompute(_d(Xe), _d(Dminv), _d(V0), _d(B))
    return Dminv.reshape(3, 3), float(V0[0]), B.reshape(9, 12)


def hex_precompute(Xe):
    Xe = f64(Xe).reshape(24)
    G, B, V0 = np.zeros(192), np.zeros(1728), np.zeros(1)
    lib().orc_hex_precompute(_d(Xe), _d(G), _d(B), _d(V0))
    return G.reshape(8, 8, 3), B.reshape(8, 9, 24), float(V0[0])


def nh_psi(F, mu, lam):
    F = f64(F).reshape(9)
    return lib().orc_nh_psi(_d(F), mu, lam)


def nh_dpsi(F, mu, lam):
    F = f64(F).reshape(9)
    P = np.zeros(9)
    lib().orc_nh_dpsi(_d(F), mu, lam, _d(P))
    return P.reshape(3, 3)


def nh_d2psi(F, mu, lam):
    F = f64(F).reshape(9)
    K = np.zeros(81)
    st = lib().orc_nh_d2psi(_d(F), mu, lam, _d(K))
    if st:
        raise RuntimeError(ERRORS.get(st, st))
    return K.reshape(9, 9)


def project_psd(K):
    K = f64(K).copy()
    n = K.shape[0]
    lib().orc_project_psd(n, _d(K.reshape(-1)))
    return K


def elem_eval(Xe, xe, E, nu, project=False, ed=None):
    Xe = f64(Xe).reshape(-1)
    xe = f64(xe).reshape(-1)
    if ed is None:
        ed = xe.size // 3
    nd = 3 * (4 if ed == TEST_POLY else ed)
    e, g, h = np.zeros(1), np.zeros(nd), np.zeros(nd * nd)
    st = lib().orc_elem_eval(ed, _d(Xe), _d(xe), E, nu, int(project), _d(e), _d(g), _d(h))
    if st:
        raise RuntimeError(ERRORS.get(st, st))
    return float(e[0]), g, h.reshape(nd, nd)


def snh_psi(F, mu, lam):
    return lib().orc_snh_psi(_d(f64(F).reshape(9)), mu, lam)


def snh_dpsi(F, mu, lam):
    P = np.zeros(9)
    lib().orc_snh_dpsi(_d(f64(F).reshape(9)), mu, lam, _d(P))
    return P.reshape(3, 3)


def snh_d2psi(F, mu, lam):
    K = np.zeros(81)
    lib().orc_snh_d2psi(_d(f64(F).reshape(9)), mu, lam, _d(K))
    return K.reshape(9, 9)


def elem_eval_ex(Xe, xe, E, nu, material=MAT_NH, muscle=None, project=False):
    """Element energy / gradient / Hessian with an elastic material and, with
    muscle = (direction[3], stiffness, activation), the SoftconMuscle energy."""
    Xe = f64(Xe).reshape(-1)
    xe = f64(xe).reshape(-1)
    nd = xe.size
    ed = nd // 3
    m, k, a = (None, 0.0, 0.0) if muscle is None else (f64(muscle[0]).reshape(3), float(muscle[1]),
                                                        float(muscle[2]))
    e, g, h = np.zeros(1), np.zeros(nd), np.zeros(nd * nd)
    st = lib().orc_elem_eval_ex(ed, _d(Xe), _d(xe), E, nu, int(material), _d(m), k, a, int(project), _d(e), _d(g),
                                _d(h))
    if st:
        raise RuntimeError(ERRORS.get(st, st))
    return float(e[0]), g, h.reshape(nd, nd)


# ---- system level --------------------------------------------------------
def cfl_timestep(q, conn, v_prev, dt, cfl=1.0):
    """Solver::calculate_CFL_timestep (solver.cpp:300-331) with
    Element::compute_minimal_edge_length (element.h:131-142): the shortest
    distance between two vertices of one element at positions q, every vertex
    pair, over all elements; the largest |v_prev| component; dt_cfl = cfl *
    length / speed when that is smaller than dt, else dt."""
    x = f64(q).reshape(-1, 3)[i32(conn)]  # E x ed x 3
    ed = x.shape[1]
    emin = min(float(np.sqrt(((x[:, i] - x[:, j]) ** 2).sum(axis=1)).min())
               for i in range(ed) for j in range(i + 1, ed))
    vmax = float(np.abs(f64(v_prev)).max())
    with np.errstate(divide="ignore", invalid="ignore"):
        dtc = cfl * np.float64(emin) / np.float64(vmax)
    return float(dtc) if dtc < dt else float(dt)


class System:
    """Energy + TimeIntegrator state of one rollout (energy.cpp:4-204)."""

    def __init__(self, X, conn, youngs, poisson, density, bcmask=None, bcvalue=None, grav=(0, 0, 0), alpha=0.0,
                 scheme=BE, method=NEWTON, ed=None, nthreads=1):
        self.X = f64(X).reshape(-1, 3)
        self.conn = i32(conn)
        nel, edc = self.conn.shape
        self.N = self.X.shape[0]
        self.n = 3 * self.N
        self.ed = edc if ed is None else ed
        self._keep = [f64(youngs), f64(poisson), f64(density),
                      i32(bcmask if bcmask is not None else np.zeros(self.n)),
                      f64(bcvalue if bcvalue is not None else np.zeros(self.n)), f64(grav)]
        Y, P, D, M, V, G = self._keep
        self.h = lib().orc_create(self.N, nel, self.ed, _d(self.X), _i(self.conn), _d(Y), _d(P), _d(D), _i(M),
                                  _d(V), _d(G), float(alpha), scheme, method)
        if not self.h:
            raise ValueError("orc_create failed")
        self.nel = nel
        lib().orc_set_threads(self.h, int(nthreads))

    @classmethod
    def from_scene(cls, sc, scheme=BE, method=NEWTON_LS, nthreads=1):
        s = cls(sc.X, sc.conn, sc.youngs, sc.poisson, sc.density, sc.bcmask, sc.bcvalue, sc.grav, sc.alpha,
                scheme=scheme, method=method, nthreads=nthreads)
        s.set_initial_velocity(sc.v0)
        return s

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().orc_free(h)
            self.h = None

    def set_threads(self, n):
        lib().orc_set_threads(self.h, int(n))

    def mass(self):
        m = np.zeros(self.n)
        lib().orc_get_mass(self.h, _d(m))
        return m

    def set_initial_velocity(self, v):
        v = f64(v).reshape(-1)
        lib().orc_set_initial_velocity(self.h, _d(v))

    def set_initial_deformation(self, q):
        q = f64(q).reshape(-1)
        lib().orc_set_initial_deformation(self.h, _d(q))

    def set_external_force(self, f):
        if f is None:
            lib().orc_set_external_force(self.h, None)
        else:
            f = f64(f).reshape(-1)
            lib().orc_set_external_force(self.h, _d(f))

    def state(self):
        q, v = np.zeros(self.n), np.zeros(self.n)
        lib().orc_get_state(self.h, _d(q), _d(v))
        return q, v

    def cfl_timestep(self, q, dt, cfl=1.0):
        """calculate_CFL_timestep at positions q with this system's v_prev."""
        return cfl_timestep(q, self.conn, self.state()[1], dt, cfl)

    def energy(self, q, dt):
        q = f64(q).reshape(-1)
        return lib().orc_energy(self.h, _d(q), float(dt))

    def gradient(self, q, dt):
        q = f64(q).reshape(-1)
        g = np.zeros(self.n)
        lib().orc_gradient(self.h, _d(q), float(dt), _d(g))
        return g

    def pattern(self, dt):
        nnz = lib().orc_pattern(self.h, float(dt))
        rp = np.zeros(self.n + 1, dtype=np.int32)
        cols = np.zeros(max(nnz, 1), dtype=np.int32)
        lib().orc_get_pattern(self.h, _i(rp), _i(cols))
        return rp, cols[:nnz]

    def scatter_map(self, dt):
        lib().orc_pattern(self.h, float(dt))
        nd = 3 * (4 if self.ed == TEST_POLY else self.ed)
        m = np.zeros(self.nel * nd * nd, dtype=np.int32)
        lib().orc_get_scatter_map(self.h, _i(m))
        return m.reshape(self.nel, nd, nd)

    def hessian(self, q, dt, project=False):
        q = f64(q).reshape(-1)
        rp, cols = self.pattern(dt)
        vals = np.zeros(max(len(cols), 1))
        st = lib().orc_hessian(self.h, _d(q), float(dt), int(project), _d(vals))
        if st:
            raise RuntimeError(ERRORS.get(st, st))
        return rp, cols, vals[: len(cols)]

    def hessian_dense(self, q, dt, project=False):
        rp, cols, vals = self.hessian(q, dt, project)
        H = np.zeros((self.n, self.n))
        for r in range(self.n):
            H[r, cols[rp[r]:rp[r + 1]]] = vals[rp[r]:rp[r + 1]]
        return H

    def step(self, q, dt, substeps=1, **kw):
        q = f64(q).reshape(-1).copy()
        o = opts(**kw)
        st = Stats()
        code = lib().orc_step(self.h, _d(q), float(dt), int(substeps), C.byref(o), C.byref(st))
        if code:
            raise RuntimeError(ERRORS.get(code, code))
        return q, st

    def minimize(self, q, dt, **kw):
        q = f64(q).reshape(-1).copy()
        o = opts(**kw)
        st = Stats()
        code = lib().orc_minimize(self.h, _d(q), float(dt), C.byref(o), C.byref(st))
        if code:
            raise RuntimeError(ERRORS.get(code, code))
        return q, st

    def bench_pcg(self, q, dt, iters, project=True):
        q = f64(q).reshape(-1)
        return lib().orc_bench_pcg(self.h, _d(q), float(dt), int(iters), int(project))

    def set_contact(self, planes=(), disks=(), surface=None, stiffness=1e5, friction=0.0, eps_v=1e-3):
        """planes: [(point, normal)], disks: [(center, normal, radius)]; surface:
        candidate node ids (None: all nodes)."""
        pp = f64([p for p, _ in planes]).reshape(-1) if planes else None
        pn = f64([n for _, n in planes]).reshape(-1) if planes else None
        dc = f64([c for c, _, _ in disks]).reshape(-1) if disks else None
        dn = f64([n for _, n, _ in disks]).reshape(-1) if disks else None
        dr = f64([r for _, _, r in disks]) if disks else None
        self._contact_keep = (pp, pn, dc, dn, dr, i32(surface) if surface is not None else None)
        s = self._contact_keep[5]
        rc = lib().orc_set_contact(self.h, len(planes), _d(pp), _d(pn), len(disks), _d(dc), _d(dn), _d(dr),
                                   0 if s is None else len(s), _i(s), float(stiffness), float(friction),
                                   float(eps_v))
        if rc != 0:
            raise ValueError(f"orc_set_contact: {rc}")

    def set_disk(self, idx, center, normal, radius):
        rc = lib().orc_set_disk(self.h, int(idx), _d(f64(center)), _d(f64(normal)), float(radius))
        if rc != 0:
            raise ValueError(f"orc_set_disk: {rc}")

    def contact_energy(self, q, dt):
        return lib().orc_contact_energy(self.h, _d(f64(q)), float(dt))

    def contact_lag(self, dt):
        lib().orc_contact_lag(self.h, float(dt))

    def set_active(self, group, fibre, sigma):
        self._active_keep = (i32(group), f64(fibre).reshape(-1))
        rc = lib().orc_set_active(self.h, int(max(group) + 1), _i(self._active_keep[0]), _d(self._active_keep[1]),
                                  float(sigma))
        if rc != 0:
            raise ValueError(f"orc_set_active: {rc}")
        self.ngroups = int(max(group) + 1)

    def set_activation(self, act):
        a = f64(act)
        assert a.size == self.ngroups
        lib().orc_set_activation(self.h, _d(a))

    def set_youngs(self, E):
        lib().orc_set_youngs(self.h, _d(f64(E)))

    def set_material(self, material):
        if lib().orc_set_material(self.h, int(material)) != 0:
            raise ValueError("orc_set_material")

    def set_muscle(self, group, stiffness, direction):
        """SoftconMuscle on elements with group[e] >= 0."""
        g = i32(group)
        self._muscle_keep = (g, f64(stiffness), f64(direction).reshape(-1))
        self.mgroups = int(g.max()) + 1
        if lib().orc_set_muscle(self.h, self.mgroups, _i(g), _d(self._muscle_keep[1]),
                                _d(self._muscle_keep[2])) != 0:
            raise ValueError("orc_set_muscle")

    def set_muscle_activation(self, act):
        a = f64(act)
        assert a.size == self.mgroups
        lib().orc_set_muscle_activation(self.h, _d(a))

    def set_pressure(self, faces, groups, ngroups=None):
        """PressureForce on faces (F x 3 for tet meshes, F x 4 for hex meshes)
        with an actuation group each; empty faces remove it."""
        f = i32(faces)
        g = i32(groups).reshape(-1)
        nv = f.shape[1] if f.ndim == 2 and f.size else (4 if self.ed == 8 else 3)
        self.pngroups = int(g.max()) + 1 if ngroups is None and g.size else int(ngroups or 0)
        self._pressure_keep = (f, g)
        if lib().orc_set_pressure(self.h, int(f.size // max(nv, 1)), int(nv), _i(f), _i(g), self.pngroups) != 0:
            raise ValueError("orc_set_pressure")

    def set_pressure_activation(self, p):
        a = f64(p).reshape(-1)
        assert a.size == self.pngroups
        lib().orc_set_pressure_activation(self.h, _d(a))

    def pressure_force(self, q):
        f = np.zeros(self.n)
        lib().orc_pressure_force(self.h, _d(f64(q).reshape(-1)), _d(f))
        return f

    def pressure_jacobian(self, q):
        """Blocks J[f, j, k] (3 x 3) of the reference's force-gradient triplets."""
        nf, nv = self._pressure_keep[0].shape
        J = np.zeros(nf * nv * nv * 9)
        lib().orc_pressure_jacobian(self.h, _d(f64(q).reshape(-1)), _d(J))
        return J.reshape(nf, nv, nv, 3, 3)

    def update_state(self, q, dt):
        q = f64(q).reshape(-1)
        lib().orc_update_state(self.h, _d(q), float(dt))

    def sqp_active(self, q):
        """PlaneContact / DiskContact update_active_constraints at q: (plane
        index per surface vertex or -1, captured disk or -1, frozen flag)."""
        ns = len(self._contact_keep[5]) if self._contact_keep[5] is not None else self.N
        pm, dm, fr = np.zeros(ns, np.int32), np.zeros(ns, np.int32), C.c_int(0)
        if lib().orc_sqp_active(self.h, _d(f64(q).reshape(-1)), _i(pm), _i(dm), C.byref(fr)) != 0:
            raise ValueError("orc_sqp_active: no contact set")
        return pm, dm, int(fr.value)


def qp_solve(rowptr, cols, P, q, vtx, a, l, u):
    """min 1/2 x^T P x + q^T x s.t. l <= A x <= u, row i of A acting on the
    DOFs of vertex vtx[i] with coefficients a[i] (OSQP's ADMM with polishing,
    run to the polished KKT point).  Returns (x, y, admm_iters)."""
    rowptr, cols, P, q = i32(rowptr), i32(cols), f64(P), f64(q)
    vtx, a, l, u = i32(vtx), f64(a).reshape(-1), f64(l), f64(u)
    n, m = len(rowptr) - 1, len(vtx)
    x, y, it = np.zeros(n), np.zeros(max(m, 1)), C.c_int(0)
    code = lib().orc_qp_solve(n, _i(rowptr), _i(cols), _d(P), _d(q), m, _i(vtx), _d(a), _d(l), _d(u), _d(x), _d(y),
                              C.byref(it))
    if code:
        raise RuntimeError(ERRORS.get(code, code))
    return x, y[:m], int(it.value)
