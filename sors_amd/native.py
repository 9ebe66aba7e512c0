"""ctypes binding of the C ABI in include/sors_amd.h (libsors_amd.so).

This is the only route from Python to the device path; there is no CPU
fallback.  Loading fails loudly when the HIP library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libsors_amd.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "sors_amd.h")

OK = 0
ERR_INVERTED, ERR_NOCONV, ERR_NAN, ERR_LINESEARCH, ERR_ARG, ERR_HIP, ERR_UNSUPPORTED = -1, -2, -3, -4, -5, -6, -7
ERR_TIMEOUT, ERR_BUDGET = -8, -9
BACKWARD_EULER, CRANK_NICOLSON = 0, 1
MINIMIZE_NEWTON, MINIMIZE_NEWTON_LINESEARCH, MINIMIZE_SQP, MINIMIZE_FROZENHESSIAN = 0, 1, 2, 3
NEOHOOKEAN, TEST_POLY, STABLE_NEOHOOKEAN = 0, 1, 2
FP64, MIXED = 0, 1
KERNEL_SPMV, KERNEL_ASSEMBLE, KERNEL_COPY, KERNEL_PCG_UPDATE = 0, 1, 2, 3

SCHEMES = {"backward_euler": BACKWARD_EULER, "crank_nicolson": CRANK_NICOLSON}
METHODS = {"minimize_newton": MINIMIZE_NEWTON, "minimize_newton_linesearch": MINIMIZE_NEWTON_LINESEARCH,
           "minimize_SQP": MINIMIZE_SQP, "minimize_frozenhessian": MINIMIZE_FROZENHESSIAN}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class SystemDesc(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int), ("n_elems", C.c_int), ("nodes_per_elem", C.c_int),
        ("rest", _dp), ("conn", _ip), ("youngs", _dp), ("poisson", _dp), ("density", _dp),
        ("material", C.c_int), ("bc_mask", _ip), ("bc_value", _dp),
        ("grav", C.c_double * 3), ("damping_alpha", C.c_double), ("scheme", C.c_int), ("method", C.c_int),
    ]


class Options(C.Structure):
    _fields_ = [
        ("precision", C.c_int), ("project_spd", C.c_int), ("max_newton", C.c_int), ("dx_tol", C.c_double),
        ("g_tol", C.c_double), ("pcg_rtol", C.c_double), ("pcg_max_iter", C.c_int), ("max_backtrack", C.c_int),
        ("pcg_check_every", C.c_int), ("use_graph", C.c_int), ("profile_kernels", C.c_int), ("predict", C.c_int),
        ("coarse_nodes", C.c_int),
        ("coarse_every", C.c_int),
        ("pcg_rtol_first", C.c_double),
        ("reuse_tangent", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("newton_iters", C.c_int), ("pcg_iters", C.c_longlong), ("backtracks", C.c_int), ("gnorm", C.c_double),
        ("dxnorm", C.c_double), ("ms_assemble", C.c_double), ("ms_solve", C.c_double),
        ("ms_linesearch", C.c_double), ("ms_total", C.c_double), ("spmv_ms", C.c_double),
        ("spmv_launches", C.c_longlong), ("assemble_kernel_ms", C.c_double), ("assemble_launches", C.c_longlong),
        ("fp64_fallbacks", C.c_int), ("qp_iters", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ContactDesc(C.Structure):
    _fields_ = [
        ("n_planes", C.c_int), ("plane_point", _dp), ("plane_normal", _dp),
        ("n_disks", C.c_int), ("disk_center", _dp), ("disk_normal", _dp), ("disk_radius", _dp),
        ("n_surface", C.c_int), ("surface_nodes", _ip),
        ("stiffness", C.c_double), ("friction", C.c_double), ("eps_v", C.c_double),
    ]


class ActiveDesc(C.Structure):
    _fields_ = [("n_groups", C.c_int), ("elem_group", _ip), ("fibre", _dp), ("sigma_max", C.c_double)]


class MuscleDesc(C.Structure):
    _fields_ = [("n_groups", C.c_int), ("elem_group", _ip), ("direction", _dp), ("stiffness", _dp)]


class PressureDesc(C.Structure):
    _fields_ = [("n_faces", C.c_int), ("faces", _ip), ("groups", _ip), ("n_groups", C.c_int)]


class ServerOptions(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_double), ("max_active", C.c_int), ("prep_threads", C.c_int),
        ("pcg_chunk", C.c_int), ("solver", Options),
    ]


class RequestStats(C.Structure):
    _fields_ = [
        ("status", C.c_int), ("steps", C.c_int), ("newton_iters", C.c_longlong), ("pcg_iters", C.c_longlong),
        ("backtracks", C.c_longlong), ("latency_ms", C.c_double), ("queue_ms", C.c_double),
        ("prep_ms", C.c_double), ("device_bytes", C.c_double), ("fp64_fallbacks", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_longlong), ("completed", C.c_longlong), ("failed", C.c_longlong),
        ("active", C.c_int), ("peak_active", C.c_int), ("resident_bytes", C.c_double),
        ("peak_resident_bytes", C.c_double), ("ticks", C.c_longlong), ("batched_launches", C.c_longlong),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SorsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


_lib = None


def lib():
    """Load libsors_amd.so; raises if it is missing (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built: run `python -m sors_amd.build` "
                              "(the HIP library is required; there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.sors_create.argtypes = [C.POINTER(SystemDesc), C.POINTER(vp)]
        L.sors_destroy.argtypes = [vp]
        L.sors_destroy.restype = None
        L.sors_last_error.argtypes = [vp]
        L.sors_last_error.restype = C.c_char_p
        L.sors_default_options.argtypes = [C.POINTER(Options)]
        L.sors_default_options.restype = None
        L.sors_set_options.argtypes = [vp, C.POINTER(Options)]
        L.sors_get_options.argtypes = [vp, C.POINTER(Options)]
        for f in ("sors_set_initial_velocity", "sors_set_initial_deformation", "sors_set_external_force",
                  "sors_get_positions"):
            getattr(L, f).argtypes = [vp, _dp]
        L.sors_get_state.argtypes = [vp, _dp, _dp]
        L.sors_get_coarse.argtypes = [vp, C.POINTER(C.c_int), _dp, _dp]
        L.sors_step.argtypes = [vp, _dp, C.c_double, C.c_int, C.POINTER(Stats)]
        L.sors_step_resident.argtypes = [vp, C.c_double, C.c_int, C.POINTER(Stats)]
        L.sors_cfl_timestep.argtypes = [vp, _dp, C.c_double, C.c_double, _dp]
        L.sors_energy.argtypes = [vp, _dp, C.c_double, _dp]
        L.sors_gradient.argtypes = [vp, _dp, C.c_double, _dp]
        L.sors_energy_noalloc.argtypes = [vp, vp, C.c_double, vp, vp]
        L.sors_pattern_nnz.argtypes = [vp, C.c_double]
        L.sors_pattern_nnz.restype = C.c_longlong
        L.sors_get_pattern.argtypes = [vp, C.c_double, _ip, _ip]
        L.sors_get_scatter_map.argtypes = [vp, C.c_double, _ip]
        L.sors_hessian.argtypes = [vp, _dp, C.c_double, C.c_int, C.c_int, _dp]
        L.sors_element_eval.argtypes = [vp, _dp, C.c_int, C.c_int, _dp, _dp, _dp]
        L.sors_linear_solve.argtypes = [vp, _dp, C.c_double, _dp, _dp, C.c_double, C.c_int, _ip]
        L.sors_bench_kernel.argtypes = [vp, C.c_int, C.c_int, _dp, _dp]
        L.sors_set_contact.argtypes = [vp, C.POINTER(ContactDesc)]
        L.sors_set_disk.argtypes = [vp, C.c_int, _dp, _dp, C.c_double]
        L.sors_set_active.argtypes = [vp, C.POINTER(ActiveDesc)]
        L.sors_set_activation.argtypes = [vp, _dp]
        L.sors_set_youngs.argtypes = [vp, _dp]
        L.sors_set_muscle.argtypes = [vp, C.POINTER(MuscleDesc)]
        L.sors_set_muscle_activation.argtypes = [vp, _dp]
        L.sors_set_pressure.argtypes = [vp, C.POINTER(PressureDesc)]
        L.sors_set_pressure_activation.argtypes = [vp, _dp]
        L.sors_update_state.argtypes = [vp, _dp, C.c_double]
        L.sors_minimize.argtypes = [vp, _dp, C.c_double, C.c_int, C.POINTER(Stats)]
        L.sors_set_state.argtypes = [vp, _dp, _dp]
        L.sors_adjoint_step.argtypes = [vp, _dp, _dp, _dp, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double,
                                        C.c_int, C.POINTER(C.c_longlong)]
        L.sors_server_default_options.argtypes = [C.POINTER(ServerOptions)]
        L.sors_server_default_options.restype = None
        L.sors_server_create.argtypes = [C.POINTER(ServerOptions), C.POINTER(vp)]
        L.sors_server_submit.argtypes = [vp, C.POINTER(SystemDesc), _dp, C.c_double, C.c_int,
                                         C.POINTER(C.c_longlong)]
        L.sors_server_wait.argtypes = [vp, C.c_longlong, _dp, C.POINTER(RequestStats), C.c_double]
        L.sors_server_get_stats.argtypes = [vp, C.POINTER(ServerStats)]
        L.sors_server_last_error.argtypes = [vp]
        L.sors_server_last_error.restype = C.c_char_p
        L.sors_server_destroy.argtypes = [vp]
        L.sors_server_destroy.restype = None
        _lib = L
    return _lib


def declared_symbols(header: str = HEADER) -> list[str]:
    """Function names declared in include/sors_amd.h."""
    text = open(header).read()
    return sorted(set(re.findall(r"^\s*[\w\s\*]+?\b(sors_\w+)\s*\(", text, flags=re.M)))


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _i(a):
    return a.ctypes.data_as(_ip) if a is not None else None


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def default_options(**kw) -> Options:
    o = Options()
    lib().sors_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Rollout:
    """One device-resident rollout: an Energy + TimeIntegrator state on the GPU."""

    def __init__(self, X, conn, youngs, poisson, density, bcmask=None, bcvalue=None, grav=(0.0, 0.0, 0.0),
                 alpha=0.0, scheme=BACKWARD_EULER, method=MINIMIZE_NEWTON_LINESEARCH, material=NEOHOOKEAN,
                 **options):
        L = lib()
        self.X = f64(X).reshape(-1, 3)
        self.conn = i32(conn)
        self.N = self.X.shape[0]
        self.n = 3 * self.N
        self.E, self.ed = self.conn.shape
        n = self.n
        self._keep = dict(Y=f64(youngs), P=f64(poisson), D=f64(density),
                          M=i32(bcmask) if bcmask is not None else None,
                          V=f64(bcvalue) if bcvalue is not None else None)
        d = SystemDesc()
        d.n_nodes, d.n_elems, d.nodes_per_elem = self.N, self.E, self.ed
        d.rest, d.conn = _d(self.X), _i(self.conn)
        d.youngs, d.poisson, d.density = _d(self._keep["Y"]), _d(self._keep["P"]), _d(self._keep["D"])
        d.material = material
        d.bc_mask = _i(self._keep["M"])
        d.bc_value = _d(self._keep["V"])
        for k in range(3):
            d.grav[k] = float(grav[k])
        d.damping_alpha = float(alpha)
        d.scheme, d.method = scheme, method
        h = C.c_void_p()
        rc = L.sors_create(C.byref(d), C.byref(h))
        if rc != OK:
            raise SorsError(rc, L.sors_last_error(None).decode())
        self.h = h
        self._n = n
        if options:
            self.set_options(**options)

    @classmethod
    def from_scene(cls, sc, method=MINIMIZE_NEWTON_LINESEARCH, scheme=BACKWARD_EULER, **options):
        r = cls(sc.X, sc.conn, sc.youngs, sc.poisson, sc.density, sc.bcmask, sc.bcvalue, sc.grav, sc.alpha,
                scheme=scheme, method=method, **options)
        r.set_initial_velocity(sc.v0)
        return r

    def close(self):
        if getattr(self, "h", None):
            lib().sors_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != OK:
            raise SorsError(rc, lib().sors_last_error(self.h).decode())

    # ---- contact -------------------------------------------------------------
    def set_contact(self, planes=(), disks=(), surface=None, stiffness=1e5, friction=0.0, eps_v=1e-3):
        """planes: [(point, normal)], disks: [(center, normal, radius)], surface:
        candidate node ids (None: every node).  None for all three removes contact."""
        if not planes and not disks and surface is None:
            self._check(lib().sors_set_contact(self.h, None))
            return
        keep = [f64([p for p, _ in planes]).reshape(-1) if planes else None,
                f64([n for _, n in planes]).reshape(-1) if planes else None,
                f64([c for c, _, _ in disks]).reshape(-1) if disks else None,
                f64([n for _, n, _ in disks]).reshape(-1) if disks else None,
                f64([r for _, _, r in disks]) if disks else None,
                i32(surface) if surface is not None else None]
        d = ContactDesc()
        d.n_planes, d.plane_point, d.plane_normal = len(planes), _d(keep[0]), _d(keep[1])
        d.n_disks, d.disk_center, d.disk_normal, d.disk_radius = len(disks), _d(keep[2]), _d(keep[3]), _d(keep[4])
        d.n_surface = 0 if keep[5] is None else len(keep[5])
        d.surface_nodes = _i(keep[5])
        d.stiffness, d.friction, d.eps_v = float(stiffness), float(friction), float(eps_v)
        self._check(lib().sors_set_contact(self.h, C.byref(d)))

    def set_disk(self, idx, center, normal, radius):
        self._check(lib().sors_set_disk(self.h, int(idx), _d(f64(center)), _d(f64(normal)), float(radius)))

    # ---- actuation / parameters / adjoint ------------------------------------
    def set_active(self, group, fibre, sigma_max):
        """Active fibre stress: group[e] in [0, n_groups) or -1, fibre [E, 3]."""
        g = i32(group)
        f = f64(fibre).reshape(-1)
        self._active_keep = (g, f)
        self.n_groups = int(g.max()) + 1
        d = ActiveDesc(self.n_groups, _i(g), _d(f), float(sigma_max))
        self._check(lib().sors_set_active(self.h, C.byref(d)))

    def set_activation(self, act):
        a = f64(act)
        assert a.size == self.n_groups
        self._check(lib().sors_set_activation(self.h, _d(a)))

    def set_muscle(self, group, stiffness, direction):
        """SoftCon muscle on elements with group[e] >= 0 (stiffness [E],
        direction [E, 3]); group None removes it."""
        if group is None:
            self._check(lib().sors_set_muscle(self.h, None))
            return
        g = i32(group)
        k = f64(stiffness)
        m = f64(direction).reshape(-1)
        self._muscle_keep = (g, k, m)
        self.muscle_groups = int(g.max()) + 1
        d = MuscleDesc(self.muscle_groups, _i(g), _d(m), _d(k))
        self._check(lib().sors_set_muscle(self.h, C.byref(d)))

    def set_muscle_activation(self, act):
        a = f64(act)
        assert a.size == self.muscle_groups
        self._check(lib().sors_set_muscle_activation(self.h, _d(a)))

    def set_pressure(self, faces, groups, n_groups=None):
        """PressureForce on faces (F x 3 on tet meshes, F x 4 on hex meshes),
        each in an actuation group; faces None removes it."""
        if faces is None:
            self._check(lib().sors_set_pressure(self.h, None))
            return
        f = i32(faces)
        g = i32(groups).reshape(-1)
        self.pressure_groups = int(n_groups) if n_groups is not None else int(g.max()) + 1
        self._pressure_keep = (f, g)
        d = PressureDesc(int(g.size), _i(f), _i(g), self.pressure_groups)
        self._check(lib().sors_set_pressure(self.h, C.byref(d)))

    def set_pressure_activation(self, p):
        a = f64(p).reshape(-1)
        assert a.size == self.pressure_groups
        self._check(lib().sors_set_pressure_activation(self.h, _d(a)))

    def minimize(self, q, dt, method=None):
        """Solver::minimize_* from the guess q (no state update)."""
        q = f64(q).reshape(-1).copy()
        st = Stats()
        m = MINIMIZE_NEWTON_LINESEARCH if method is None else method
        self._check(lib().sors_minimize(self.h, _d(q), float(dt), int(m), C.byref(st)))
        return q, st

    def update_state(self, q, dt):
        self._check(lib().sors_update_state(self.h, _d(f64(q).reshape(-1)), float(dt)))

    def set_youngs(self, E):
        e = f64(E)
        assert e.size == self.E
        self._check(lib().sors_set_youngs(self.h, _d(e)))

    def set_state(self, q=None, v=None):
        qq = f64(q).reshape(-1) if q is not None else None
        vv = f64(v).reshape(-1) if v is not None else None
        self._check(lib().sors_set_state(self.h, _d(qq), _d(vv)))

    def adjoint_step(self, q_prev, v_prev, q_next, dt, qbar_next, vbar_next, youngs_bar=None, act_bar=None,
                     rtol=1e-10, max_iter=20000):
        """One backward-Euler step's adjoint: returns (qbar_prev, vbar_prev, pcg_iters);
        youngs_bar / act_bar (float64 arrays) are accumulated in place."""
        qb, vb = np.empty(self.n), np.empty(self.n)
        it = C.c_longlong()
        args = [f64(q_prev).reshape(-1), f64(v_prev).reshape(-1), f64(q_next).reshape(-1),
                f64(qbar_next).reshape(-1), f64(vbar_next).reshape(-1)]
        for a in (youngs_bar, act_bar):
            assert a is None or (a.dtype == np.float64 and a.flags.c_contiguous)
        self._check(lib().sors_adjoint_step(self.h, _d(args[0]), _d(args[1]), _d(args[2]), float(dt), _d(args[3]),
                                            _d(args[4]), _d(qb), _d(vb), _d(youngs_bar), _d(act_bar), float(rtol),
                                            int(max_iter), C.byref(it)))
        return qb, vb, it.value

    # ---- options / state -------------------------------------------------
    def options(self) -> Options:
        o = Options()
        self._check(lib().sors_get_options(self.h, C.byref(o)))
        return o

    def set_options(self, **kw):
        o = self.options()
        for k, v in kw.items():
            setattr(o, k, v)
        self._check(lib().sors_set_options(self.h, C.byref(o)))

    def set_initial_velocity(self, v):
        v = f64(v).reshape(-1)
        self._check(lib().sors_set_initial_velocity(self.h, _d(v)))

    def set_initial_deformation(self, q):
        q = f64(q).reshape(-1)
        self._check(lib().sors_set_initial_deformation(self.h, _d(q)))

    def set_external_force(self, f):
        if f is None:
            self._check(lib().sors_set_external_force(self.h, None))
        else:
            f = f64(f).reshape(-1)
            self._check(lib().sors_set_external_force(self.h, _d(f)))

    def state(self):
        q, v = np.zeros(self.n), np.zeros(self.n)
        self._check(lib().sors_get_state(self.h, _d(q), _d(v)))
        return q, v

    def positions(self):
        q = np.zeros(self.n)
        self._check(lib().sors_get_positions(self.h, _d(q)))
        return q

    def coarse_operator(self):
        """(E, E^-1) of the two-level preconditioner: the last refresh's
        Galerkin matrix and the inverse in use (None when no coarse
        correction is active)."""
        nc = C.c_int(0)
        self._check(lib().sors_get_coarse(self.h, C.byref(nc), None, None))
        if nc.value == 0:
            return None
        E = np.zeros((nc.value, nc.value))
        Einv = np.zeros_like(E)
        self._check(lib().sors_get_coarse(self.h, C.byref(nc), _d(E), _d(Einv)))
        return E, Einv

    def cfl_timestep(self, q, dt, cfl=1.0):
        """Solver::calculate_CFL_timestep at positions q (None: the resident
        ones) with this rollout's v_prev."""
        out = np.zeros(1)
        qa = None if q is None else f64(q).reshape(-1)
        self._check(lib().sors_cfl_timestep(self.h, _d(qa), float(dt), float(cfl), _d(out)))
        return float(out[0])

    # ---- stepping --------------------------------------------------------
    def step(self, q, dt, substeps=1):
        q = f64(q).reshape(-1).copy()
        st = Stats()
        self._check(lib().sors_step(self.h, _d(q), float(dt), int(substeps), C.byref(st)))
        return q, st

    def step_resident(self, dt, substeps=1):
        st = Stats()
        self._check(lib().sors_step_resident(self.h, float(dt), int(substeps), C.byref(st)))
        return st

    # ---- inspection ------------------------------------------------------
    def energy(self, q, dt):
        q = f64(q).reshape(-1)
        out = np.zeros(1)
        self._check(lib().sors_energy(self.h, _d(q), float(dt), _d(out)))
        return float(out[0])

    def gradient(self, q, dt):
        q = f64(q).reshape(-1)
        g = np.zeros(self.n)
        self._check(lib().sors_gradient(self.h, _d(q), float(dt), _d(g)))
        return g

    def pattern(self, dt):
        nnz = lib().sors_pattern_nnz(self.h, float(dt))
        if nnz < 0:
            self._check(int(nnz))
        rp = np.zeros(self.n + 1, np.int32)
        cols = np.zeros(max(nnz, 1), np.int32)
        self._check(lib().sors_get_pattern(self.h, float(dt), _i(rp), _i(cols)))
        return rp, cols[:nnz]

    def scatter_map(self, dt):
        nd = 3 * self.ed
        m = np.zeros(self.E * nd * nd, np.int32)
        self._check(lib().sors_get_scatter_map(self.h, float(dt), _i(m)))
        return m.reshape(self.E, nd, nd)

    def hessian(self, q, dt, precision=FP64, project=False):
        q = f64(q).reshape(-1)
        rp, cols = self.pattern(dt)
        vals = np.zeros(max(len(cols), 1))
        self._check(lib().sors_hessian(self.h, _d(q), float(dt), int(precision), int(project), _d(vals)))
        return rp, cols, vals[: len(cols)]

    def hessian_dense(self, q, dt, precision=FP64, project=False):
        rp, cols, vals = self.hessian(q, dt, precision, project)
        H = np.zeros((self.n, self.n))
        for r in range(self.n):
            H[r, cols[rp[r]:rp[r + 1]]] = vals[rp[r]:rp[r + 1]]
        return H

    def element_eval(self, q, precision=FP64, project=False):
        q = f64(q).reshape(-1)
        nd = 3 * self.ed
        e, g, h = np.zeros(self.E), np.zeros(self.E * nd), np.zeros(self.E * nd * nd)
        self._check(lib().sors_element_eval(self.h, _d(q), int(precision), int(project), _d(e), _d(g), _d(h)))
        return e, g.reshape(self.E, nd), h.reshape(self.E, nd, nd)

    def bench_kernel(self, kernel, reps=20):
        ms, by = np.zeros(1), np.zeros(1)
        self._check(lib().sors_bench_kernel(self.h, int(kernel), int(reps), _d(ms), _d(by)))
        return float(ms[0]), float(by[0])

    def linear_solve(self, q, dt, b, rtol=1e-10, max_iter=20000):
        q = f64(q).reshape(-1)
        b = f64(b).reshape(-1)
        x = np.zeros(self.n)
        it = C.c_int(0)
        self._check(lib().sors_linear_solve(self.h, _d(q), float(dt), _d(b), _d(x), float(rtol), int(max_iter),
                                            C.byref(it)))
        return x, it.value


def _scene_desc(sc, method, scheme):
    """SystemDesc over a Scene's arrays; returns (desc, keep-alive tuple)."""
    X, conn = f64(sc.X).reshape(-1, 3), i32(sc.conn)
    keep = (X, conn, f64(sc.youngs), f64(sc.poisson), f64(sc.density),
            i32(sc.bcmask) if sc.bcmask is not None else None,
            f64(sc.bcvalue) if sc.bcvalue is not None else None)
    d = SystemDesc()
    d.n_nodes, d.n_elems, d.nodes_per_elem = X.shape[0], conn.shape[0], conn.shape[1]
    d.rest, d.conn = _d(keep[0]), _i(keep[1])
    d.youngs, d.poisson, d.density = _d(keep[2]), _d(keep[3]), _d(keep[4])
    d.material = NEOHOOKEAN
    d.bc_mask, d.bc_value = _i(keep[5]), _d(keep[6])
    for k in range(3):
        d.grav[k] = float(sc.grav[k])
    d.damping_alpha = float(sc.alpha)
    d.scheme, d.method = scheme, method
    return d, keep


class Server:
    """Serving loop: continuous batching of independent rollouts on one GPU.

    submit() queues a rollout (mesh, material, initial velocity, steps) and
    returns its id at once; wait(id) returns its final positions and stats.
    """

    def __init__(self, hbm_budget_bytes=200e9, max_active=64, prep_threads=4, pcg_chunk=16, **solver):
        L = lib()
        o = ServerOptions()
        L.sors_server_default_options(C.byref(o))
        o.hbm_budget_bytes = float(hbm_budget_bytes)
        o.max_active, o.prep_threads, o.pcg_chunk = max_active, prep_threads, pcg_chunk
        for k, v in solver.items():
            setattr(o.solver, k, v)
        h = C.c_void_p()
        rc = L.sors_server_create(C.byref(o), C.byref(h))
        if rc != OK:
            raise SorsError(rc, L.sors_server_last_error(None).decode())
        self.h = h
        self._n = {}

    def _check(self, rc):
        if rc != OK:
            raise SorsError(rc, lib().sors_server_last_error(self.h).decode())

    def submit(self, sc, steps, dt, method=MINIMIZE_NEWTON_LINESEARCH, scheme=BACKWARD_EULER) -> int:
        d, keep = _scene_desc(sc, method, scheme)
        v0 = f64(sc.v0).reshape(-1) if sc.v0 is not None else None
        rid = C.c_longlong()
        self._check(lib().sors_server_submit(self.h, C.byref(d), _d(v0), float(dt), int(steps), C.byref(rid)))
        self._n[rid.value] = 3 * d.n_nodes
        return rid.value

    def wait(self, rid, timeout=-1.0):
        """(q [N,3], stats dict); raises SorsError with the request's status on failure,
        ERR_TIMEOUT if it is not finished within `timeout` seconds."""
        q = np.empty(self._n[rid], dtype=np.float64)
        st = RequestStats()
        rc = lib().sors_server_wait(self.h, rid, _d(q), C.byref(st), float(timeout))
        if rc == ERR_TIMEOUT:
            raise SorsError(rc, "request not finished")
        self._n.pop(rid, None)
        if rc != OK:
            raise SorsError(rc, lib().sors_server_last_error(self.h).decode())
        return q.reshape(-1, 3), st.as_dict()

    def stats(self):
        st = ServerStats()
        self._check(lib().sors_server_get_stats(self.h, C.byref(st)))
        return st.as_dict()

    def close(self):
        if getattr(self, "h", None):
            lib().sors_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
