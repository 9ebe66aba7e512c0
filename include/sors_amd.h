/*
 * sors_amd.h -- C ABI of the MI355X implicit Newton step for energy-based
 * FEM (Neo-Hookean tet4 / hex8), a drop-in for the CPU/Eigen path of SORS.
 *
 * Plain C types only: host pointers and sizes in, status codes out.  Every
 * entry point returns SORS_OK (0) or a negative SORS_ERR_* code; the message
 * is available from sors_last_error(ctx) (or sors_last_error(NULL) for
 * sors_create).  Arrays are node-major FP64 like the reference's state
 * vector q = [x0,y0,z0,x1,...] (cpp/src/energy.cpp:216-221); element-local
 * DOF = 3*localNode + comp.
 *
 * Threading: one context per caller thread at a time; calls on one context
 * must be serialised.  Each context owns its device buffers and one HIP
 * stream.  Host pointers are borrowed for the duration of a call.
 */
#ifndef SORS_AMD_H
#define SORS_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define SORS_OK 0
#define SORS_ERR_INVERTED (-1)   /* "Inverted Elements" (cpp/src/energies/neoHookean.cpp:97-101) */
#define SORS_ERR_NOCONV (-2)     /* "Newton's method did not converge" (cpp/src/solver.cpp:100-103) */
#define SORS_ERR_NAN (-3)        /* step has NaN (cpp/src/solver.cpp:90-93) */
#define SORS_ERR_LINESEARCH (-4) /* Armijo backtracking cap reached */
#define SORS_ERR_ARG (-5)
#define SORS_ERR_HIP (-6)
#define SORS_ERR_UNSUPPORTED (-7)
#define SORS_ERR_TIMEOUT (-8)     /* sors_server_wait: request not finished yet */
#define SORS_ERR_BUDGET (-9)      /* request can never fit the server's HBM budget */

/* SimulationSettings::timeSteppingScheme (cpp/headers/simulationSettings.h:31) */
enum { SORS_BACKWARD_EULER = 0, SORS_CRANK_NICOLSON = 1 };
/* SimulationSettings::solverMethod (cpp/src/solver.cpp:28-40).  The line
 * search variant keeps minimize_newton's Dirichlet handling.
 * SQP (minimize_SQP, solver.cpp:109-197): the unmasked assembly semantics
 * (energy.cpp:275,337) with the Dirichlet mask as NeumannBC equality rows
 * x - xPrev - dt v = 0 (neumannBC.cpp:4-40) and the planes and disks of
 * sors_set_contact as the PlaneContact / DiskContact inequality rows
 * n.(x - p) >= 0 on their active surface vertices (planeContact.cpp,
 * diskContact.cpp; active sets updated once per step from the initial guess,
 * the disk footprint frozen between steps as the reference does); each SQP
 * iteration solves its QP exactly (working-set iterations of projected PCG
 * solves, where the reference calls OSQP) and takes the full step, until
 * ||dx|| < dx_tol; no predictor (the caller's guess is the start).
 * FROZENHESSIAN (minimize_frozenhessian, solver.cpp:268-294): the tangent of
 * the step's initial guess, and its preconditioner, serve every iteration;
 * each iteration assembles the gradient alone; no line search; stopping after
 * max_newton iterations without ||dx|| < dx_tol is not an error (the
 * reference prints a message and returns the iterate). */
enum {
    SORS_MINIMIZE_NEWTON = 0,
    SORS_MINIMIZE_NEWTON_LINESEARCH = 1,
    SORS_MINIMIZE_SQP = 2,
    SORS_MINIMIZE_FROZENHESSIAN = 3
};
/* ElementEnergy plugins (cpp/headers/energies/elementEnergy.h): compressible
 * Neo-Hookean ("neohookean", cpp/src/energies/neoHookean.cpp), the polynomial
 * test energy of the reference's unit tests (cpp/src/_test.cpp:8-66), and
 * Stable Neo-Hookean ("stableneohookean", cpp/src/energies/stableNeoHookean.cpp,
 * with its parameters mu' = 4/3 mu, lambda' = lambda + 5/8 mu',
 * alpha = 1 + 3/4 mu'/lambda' from each element's youngsModulus and
 * poissonsRatio).  The SoftCon muscle energy is added per element with
 * sors_set_muscle. */
enum { SORS_NEOHOOKEAN = 0, SORS_TEST_POLY = 1, SORS_STABLE_NEOHOOKEAN = 2 };
enum { SORS_FP64 = 0, SORS_MIXED = 1 };

/* Arguments of Energy<3,E>::Energy (cpp/headers/energy.h:73-84) as arrays:
 * undeformedVertices, eleIdx, per-element Params "density" /
 * "youngsModulus" / "poissonsRatio", the "neumannBC" constraint's
 * neumannBCmask / neumannBCvalue (energy.cpp:61-68), gravAcceleration,
 * dampingAlpha and the SimulationSettings fields that Energy reads. */
typedef struct {
    int n_nodes;
    int n_elems;
    int nodes_per_elem;     /* 4 (Tetrahedron) or 8 (Hexahedron) */
    const double* rest;     /* [n_nodes][3] */
    const int* conn;        /* [n_elems][nodes_per_elem] */
    const double* youngs;   /* [n_elems] */
    const double* poisson;  /* [n_elems] */
    const double* density;  /* [n_elems] */
    int material;           /* SORS_NEOHOOKEAN | SORS_TEST_POLY | SORS_STABLE_NEOHOOKEAN */
    const int* bc_mask;     /* [3 n_nodes] or NULL */
    const double* bc_value; /* [3 n_nodes] prescribed velocity or NULL */
    double grav[3];
    double damping_alpha;
    int scheme;             /* SORS_BACKWARD_EULER | SORS_CRANK_NICOLSON */
    int method;             /* SORS_MINIMIZE_NEWTON[_LINESEARCH] | _SQP | _FROZENHESSIAN */
} sors_system_desc;

/* Solver controls.  Defaults (sors_default_options): mixed precision, SPD
 * projection on, MAX_ITER = 100 and DX_TOL = 1e-9 (cpp/headers/common.h:21-25),
 * Armijo alpha0 = 1, beta = 0.5, c = 0.1 (cpp/src/solver.cpp:208-210). */
typedef struct {
    int precision;        /* SORS_FP64: all-FP64; SORS_MIXED: FP32 element tangents,
                             CSR values, SpMV and PCG vectors; FP64 state, residual,
                             energies and every dot product */
    int project_spd;      /* clamp each element tangent to PSD (needed by PCG) */
    int max_newton;
    double dx_tol;        /* converged when ||accepted step||_2 < dx_tol */
    double g_tol;         /* also converged when ||g||_2 < g_tol (<= 0: off) */
    double pcg_rtol;      /* relative residual of each linear solve */
    int pcg_max_iter;
    int max_backtrack;
    int pcg_check_every;  /* PCG iterations before the first host convergence check; later
                             checks come after 1/4, 1/2 or all of that many, the fewest
                             expected to end the solve at its convergence rate so far */
    int use_graph;        /* replay PCG iteration chunks from a captured hipGraph */
    int profile_kernels;  /* time SpMV and assembly launches with HIP events */
    int predict;          /* Newton's start instead of the caller's guess: 0 the
                             guess, 1 qPrev + dt*vPrev, 2 (default) qPrev + dt*(vPrev
                             + (vPrev - v before the last step)) once a step has been
                             taken since the state was set; Dirichlet DOFs start at
                             qPrev + dt*value.  The converged state is the same
                             minimiser */
    int coarse_nodes;     /* > 0 (default 2048): two-level PCG preconditioner,
                             block-Jacobi plus a coarse correction on the
                             rigid-body modes of spatial node aggregates of
                             about this many nodes (sliced node-block layout
                             only); 0: block-Jacobi alone */
    int coarse_every;     /* coarse operator refreshed on every k-th tangent
                             (default 16; <= 1: every tangent, inverted
                             before the solve); with k > 1 a refresh is
                             inverted on a second stream ahead of the
                             solve's kernels (the host does not wait), the
                             solve runs with the previous inverse, and the
                             next tangent installs it.  PCG always runs
                             with an SPD preconditioner */
    double pcg_rtol_first; /* > pcg_rtol: relative PCG tolerance of the first
                             Newton solve of each minimisation (default 0: off,
                             every solve uses pcg_rtol).  The first Newton step
                             from the predicted start is limited by the
                             problem's nonlinearity (its residual falls by about
                             1e-2 however exactly it is solved), so a looser
                             first solve saves PCG iterations without changing
                             the converged state, which g_tol / dx_tol fix */
    int reuse_tangent;    /* 1: when the last Newton solve's forcing term predicts a
                             residual within reach of g_tol (the next solve would
                             be aimed at it), the next iteration evaluates the
                             gradient alone first and, if the residual is within
                             reach but not converged, solves with the last
                             tangent and its preconditioner once more instead of
                             assembling a new one (default 0: every solve gets a
                             fresh tangent).  Needs g_tol > 0; the converged state
                             is fixed by g_tol / dx_tol as before */
} sors_options;

typedef struct {
    int newton_iters;
    long long pcg_iters;
    int backtracks;
    double gnorm;         /* ||g||_2 at the last evaluated iterate (FP64) */
    double dxnorm;        /* ||last accepted step||_2 */
    double ms_assemble;
    double ms_solve;
    double ms_linesearch;
    double ms_total;
    double spmv_ms;            /* summed SpMV kernel time (profile_kernels) */
    long long spmv_launches;
    double assemble_kernel_ms; /* summed element-assembly kernel time of the tangent (gradient +
                                  Hessian) assemblies (profile_kernels) */
    long long assemble_launches; /* tangent assemblies timed; gradient-only residual checks are not */
    int fp64_fallbacks;   /* mixed-precision Newton solves that did not converge within max_newton
                             and were continued from their iterate in FP64 (see sors_step) */
    int qp_iters;         /* minimize_SQP: working-set iterations (one projected PCG solve each) */
} sors_stats;

typedef struct sors_ctx sors_ctx;

/* Energy constructor + TimeIntegrator(qPrev = rest, vPrev = 0)
 * (cpp/src/energy.cpp:4-204, cpp/headers/integrator.h:155-161). */
int sors_create(const sors_system_desc* desc, sors_ctx** out);
void sors_destroy(sors_ctx* ctx);
const char* sors_last_error(const sors_ctx* ctx);
void sors_default_options(sors_options* opt);
int sors_set_options(sors_ctx* ctx, const sors_options* opt);
int sors_get_options(const sors_ctx* ctx, sors_options* opt);

/* Energy::set_initial_velocity / set_initial_deformation (cpp/headers/energy.h:203-204) */
int sors_set_initial_velocity(sors_ctx* ctx, const double* v);
int sors_set_initial_deformation(sors_ctx* ctx, const double* q);
/* "vertexForce" external force (cpp/src/forces/vertexForce.cpp:4-8): the
 * actuation vector [3 n_nodes]; NULL removes it. */
int sors_set_external_force(sors_ctx* ctx, const double* f);
int sors_get_state(sors_ctx* ctx, double* q_prev, double* v_prev);

/* Solver::step (cpp/src/solver.cpp:5-44): q is the initial guess on entry
 * and the new positions on return; dt == 0 is rejected (the caller resolves
 * settings.dt); substeps >= 1.  stats may be NULL.  In mixed precision, a
 * Newton solve that does not converge within max_newton iterations goes on
 * from its iterate with FP64 tangents and solves (another max_newton at
 * most) before "Newton's method did not converge" is returned. */
int sors_step(sors_ctx* ctx, double* q, double dt, int substeps, sors_stats* stats);
/* Same, starting from the device-resident state of the previous step (no
 * host transfer); read positions with sors_get_positions. */
int sors_step_resident(sors_ctx* ctx, double dt, int substeps, sors_stats* stats);
int sors_get_positions(sors_ctx* ctx, double* q);
/* Solver::calculate_CFL_timestep (cpp/src/solver.cpp:300-331): dt_cfl =
 * cfl * (shortest distance between two vertices of one element at positions
 * q, every vertex pair as in cpp/headers/geometries/element.h:131-142) /
 * max |v_prev| component, or dt when that is not smaller.  q == NULL uses the
 * resident positions.  Solver::step (solver.cpp:16-20) then takes
 * ceil(dt / dt_cfl) substeps. */
int sors_cfl_timestep(sors_ctx* ctx, const double* q, double dt, double cfl, double* dt_cfl);

/* ---- inspection entry points used by the parity tests ---------------- */
/* Energy::compute_energy (cpp/src/energy.cpp:208-248) */
int sors_energy(sors_ctx* ctx, const double* q, double dt, double* energy);
/* Allocation-free variant of sors_energy.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * q_dev [3 n_nodes] and energy_dev [1] are device pointers; the result is
 * written to energy_dev by kernels enqueued on `stream` (a hipStream_t;
 * NULL: the context's stream), and the call returns without waiting: read
 * energy_dev after synchronising that stream.  Same kernels and summation
 * order as sors_energy, so the value is identical.  The reduction partials
 * live in one static device array per GPU (16,448 bytes, held until the
 * process ends), not in the context: the call allocates nothing, and every
 * call on one GPU shares that array, so calls that overlap on one GPU (two
 * streams, two threads, two contexts) race and must be serialised by the
 * caller.  The context's own buffers are read, not written. */
int sors_energy_noalloc(sors_ctx* ctx, const double* q_dev, double dt, double* energy_dev, void* stream);
/* Energy::compute_gradient (cpp/src/energy.cpp:252-308), FP64 */
int sors_gradient(sors_ctx* ctx, const double* q, double dt, double* g);
/* CSR pattern of Energy::compute_hessian for this dt (pattern rule in the
 * library's pattern.h); rowptr[3 n_nodes + 1], cols[nnz] */
long long sors_pattern_nnz(sors_ctx* ctx, double dt);
int sors_get_pattern(sors_ctx* ctx, double dt, int* rowptr, int* cols);
/* scatter map [n_elems][(3 ed)^2]: CSR value index or -1 (energy.cpp:337) */
int sors_get_scatter_map(sors_ctx* ctx, double dt, int* map);
/* assembled Hessian values on the pattern, returned as FP64 */
int sors_hessian(sors_ctx* ctx, const double* q, double dt, int precision, int project, double* vals);
/* per-element energy [E], gradient [E][3ed], Hessian [E][3ed][3ed] (any may be NULL);
 * Element::compute_energy / Tetrahedron::compute_gradient,compute_hessian
 * (cpp/headers/geometries/element.h:40-58, cpp/src/geometries/tetrahedron.cpp:99-134,
 * cpp/src/geometries/hexahedron.cpp:68-103) */
int sors_element_eval(sors_ctx* ctx, const double* q, int precision, int project, double* energy, double* grad,
                      double* hess);
/* solve H(q) x = b with the device block-Jacobi PCG (precision per options) */
int sors_linear_solve(sors_ctx* ctx, const double* q, double dt, const double* b, double* x, double rtol,
                      int max_iter, int* iters);

/* The two-level preconditioner's coarse operator (no reference counterpart:
 * it replaces the reference's direct factorisation, solver.cpp:82-84, by a
 * PCG whose coarse correction needs E = Z^T A Z and E^-1).  nc receives the
 * coarse size (0 when no coarse correction is in use); E (nc x nc, the last
 * refresh's Galerkin matrix, row-major) and Einv (the inverse in use) are
 * copied when not NULL.  With coarse_every = 1 the two belong together. */
int sors_get_coarse(sors_ctx* ctx, int* nc, double* E, double* Einv);

/* ---- kernel micro-benchmarks (profiling runs) ------------------------ */
enum { SORS_KERNEL_SPMV = 0, SORS_KERNEL_ASSEMBLE = 1, SORS_KERNEL_COPY = 2, SORS_KERNEL_PCG_UPDATE = 3 };
/* Launch one kernel `reps` times back to back on the context's stream, on the
 * matrix / state left by the last step, timed with HIP events.  avg_ms is
 * the mean duration of one launch; algo_bytes the compulsory HBM bytes of
 * one launch (SpMV: vals + cols + rowptr + x + y; COPY: n values read and
 * written; ASSEMBLE: connectivity, Lame constants, node positions, CSR
 * values and gradient written once). */
int sors_bench_kernel(sors_ctx* ctx, int kernel, int reps, double* avg_ms, double* algo_bytes);

/* ---- Contact (penalty + lagged Coulomb friction) ---------------------------
 * The reference resolves contact as hard constraints n.(x - p) >= 0 by SQP
 * over OSQP (cpp/src/solver.cpp:109-197) with its "planeContact" and
 * "diskContact" constraints (cpp/src/energy.cpp:69-105).  This device path
 * keeps their geometry -- planes (point, normal; Plane3D, plane3D.h:30-33)
 * and disks (centre, normal, radius; Disk3D, disk3D.h:33-45) acting on the
 * surface vertices -- and adds to the step objective, per candidate node:
 *   penalty   1/2 k w d^2 for d = n.(x - p) < 0 (w = 1 for a plane; for a
 *             disk w = 1 inside the radius, a C1 cubic falloff over 1 mm
 *             outside it, 0 beyond);
 *   friction  mu N phi(|uT|): N, n and the primitive's displacement D over
 *             the step are taken at the start of each step (largest penalty
 *             force at qprev); u = x - qprev - D, uT its tangential part,
 *             phi(y) = y^2/(2 eps) below eps = eps_v dt, y - eps/2 above.
 * Parameters keep the reference's names: positionVectorListPlanes /
 * normalVectorListPlanes, diskCenterVectorList / diskNormalVectorList /
 * diskRadiusList, surfaceVerticesIdx.  Normals are normalised.  Applies to
 * minimize_newton(_linesearch) and the frozen Hessian; masked DOFs take no
 * contact terms.  Under SORS_MINIMIZE_SQP the same primitives are the
 * reference's hard constraints instead (see the solverMethod enum) and
 * stiffness, friction and eps_v are not used. */
typedef struct {
    int n_planes;                   /* <= 4 */
    const double* plane_point;      /* [n_planes][3] */
    const double* plane_normal;     /* [n_planes][3] */
    int n_disks;                    /* <= 4 */
    const double* disk_center;      /* [n_disks][3] */
    const double* disk_normal;      /* [n_disks][3] */
    const double* disk_radius;      /* [n_disks] */
    int n_surface;                  /* candidate nodes; 0 with NULL: every node */
    const int* surface_nodes;
    double stiffness;               /* k (N/m per node) */
    double friction;                /* Coulomb mu (0: frictionless) */
    double eps_v;                   /* stick/slip velocity threshold (m/s) */
} sors_contact_desc;

/* Set (or with NULL remove) contact; recomputes the friction lag from the
 * current state.  Replaces the planeContact / diskContact entries of
 * Energy::Energy's constraint list (energy.cpp:69-105). */
int sors_set_contact(sors_ctx* ctx, const sors_contact_desc* desc);
/* DiskContact::set_disk (cpp/headers/constraints/diskContact.h:72-80): move
 * disk idx; the displacement since the last completed step drives the
 * friction of the next step. */
int sors_set_disk(sors_ctx* ctx, int idx, const double* center, const double* normal, double radius);

/* ---- Actuation and the adjoint -------------------------------------------
 * Active fibre stress (tet4): Psi_act = 1/2 sigma_max a (|F f|^2 - 1) per
 * unit rest volume, with the element's unit fibre direction f and the
 * activation a of its muscle group (first Piola stress sigma_max a (F f) f^T;
 * at rest the active stress sigma_max a f (x) f).  It plays the role of the
 * reference's actuated element energies, whose activations arrive per
 * muscle group through the "actuation" Params of Solver::step (SoftCon
 * muscle, cpp/src/energies/softconMuscle.cpp:12-134; element.h:52).
 * elem_group[e] = -1 leaves element e passive. */
typedef struct {
    int n_groups;
    const int* elem_group;     /* [n_elems] group in [0, n_groups) or -1 */
    const double* fibre;       /* [n_elems][3] (normalised) */
    double sigma_max;          /* Pa */
} sors_active_desc;

int sors_set_active(sors_ctx* ctx, const sors_active_desc* desc);   /* NULL removes it */
/* activations [n_groups] used by the following steps (Solver::step's actuation) */
int sors_set_activation(sors_ctx* ctx, const double* act);

/* SoftconMuscleEnergy (cpp/src/energies/softconMuscle.cpp:12-134) on the
 * elements with elem_group[e] >= 0, with their "muscleGroup",
 * "softconStiffness" and "softconDirection" Params
 * (cpp/src/geometries/tetrahedron.cpp:82-88, hexahedron.cpp:51-57); tet4 and
 * hex8.  Per Gauss point, u = F m, l = |u|, L = 1 - activation, r = L / l:
 * energy k/2 |l - L| (as the reference computes it, :27-31), gradient
 * k (1 - r) u m^T (:55-72, the gradient of k/2 (l - L)^2), Hessian
 * k [(r/l^2) u u^T + (1 - r) 1 1^T] (x) m m^T (:98-130); with project_spd
 * its 3x3 factor is clamped to PSD.  NULL removes it. */
typedef struct {
    int n_groups;
    const int* elem_group;     /* [n_elems] group in [0, n_groups) or -1 */
    const double* direction;   /* [n_elems][3] m, as given (not normalised) */
    const double* stiffness;   /* [n_elems] k */
} sors_muscle_desc;

int sors_set_muscle(sors_ctx* ctx, const sors_muscle_desc* desc);
/* the "softconMuscle" actuation of Solver::step (element.h:52): activations
 * [n_groups] used by the following steps (0 until set) */
int sors_set_muscle_activation(sors_ctx* ctx, const double* act);
/* PressureForce (cpp/src/forces/pressureForce.cpp:14-185; the "pressure"
 * force of Energy, energy.cpp:155-162 with forceParameterList
 * "surfaceVertexIdx" / "surfaceGroups"): faces of 3 vertices (tet4 meshes) or
 * 4 (hex8), faces[nv f + j], each in an actuation group; the force -p A / nv
 * on each face vertex (A: the face's area vector at the current positions).
 * It enters the energy as +f(q).(q - qPrev) and the gradient as -f(q)
 * (energy.cpp:228-244, 282-293).  Its Jacobian (the reference's p/6 blocks,
 * minimize_newton only) is part of sors_hessian's result; the solve's SPD
 * tangent leaves it out, so Newton's iterates differ from the reference's
 * but not its converged state.  Needs the node-block layout; n_faces = 0 or
 * NULL removes the force.  Pressures start at 0. */
typedef struct {
    int n_faces;
    const int* faces;         /* [n_faces][3 or 4] vertex indices */
    const int* groups;        /* [n_faces] actuation group */
    int n_groups;
} sors_pressure_desc;
int sors_set_pressure(sors_ctx* ctx, const sors_pressure_desc* desc);
/* the "pressure" actuation: pressure per group [n_groups] */
int sors_set_pressure_activation(sors_ctx* ctx, const double* pressure);
/* per-element Young's modulus (Lame constants recomputed with each element's
 * Poisson ratio, neoHookean.cpp:10-11) */
int sors_set_youngs(sors_ctx* ctx, const double* youngs);
/* Energy::update_state (cpp/src/energy.cpp:539-557, integrator.h:79-83): q_prev
 * = q, v_prev from q by the integrator's rule (BE: (q - q_prev)/dt; CN:
 * 2 (q - q_prev)/dt - v_prev), g_prev = the internal force of the last
 * gradient evaluation (gPrevTmp_).  The positions become q. */
int sors_update_state(sors_ctx* ctx, const double* q, double dt);
/* Solver::minimize_newton / minimize_newton_linesearch / minimize_frozenhessian
 * (cpp/src/solver.cpp:56-107, 201-266, 268-298): the minimiser of the step
 * energy from the initial guess q, returned in q; the integrator state is not
 * updated (Solver::step = minimize + update_state).  A mixed-precision solve
 * that reaches max_newton goes on in FP64, as in sors_step.  minimize_SQP
 * (method SORS_MINIMIZE_SQP) updates the contact active sets from q first,
 * as the reference's minimize_SQP does. */
int sors_minimize(sors_ctx* ctx, double* q, double dt, int method, sors_stats* stats);
/* TimeIntegrator state (integrator.h:101-102): qPrev = q (and the positions),
 * vPrev = v; either may be NULL to keep it */
int sors_set_state(sors_ctx* ctx, const double* q, const double* v);

/* Adjoint of one backward Euler step (no reference counterpart: the
 * reference has no gradients).  Given the state before the step (q_prev,
 * v_prev), the converged positions after it (q_next, v_next = (q_next -
 * q_prev)/dt), the current activations and Young's moduli, and the
 * adjoints qbar_next = dL/dq_next, vbar_next = dL/dv_next, it solves
 *   H lambda = qbar_next + vbar_next/dt     (H: the unprojected FP64 Hessian
 *                                            of the step at q_next, PCG)
 * and returns qbar_prev = dL/dq_prev, vbar_prev = dL/dv_prev (overwritten)
 * and ADDS dL/dE per element into youngs_bar and dL/da per group into
 * act_bar (either may be NULL).  Backward Euler, minimize_newton(_linesearch),
 * tet4; frictionless plane contact is included, friction and disks are not
 * (SORS_ERR_UNSUPPORTED).  Overwrites the context's state. */
int sors_adjoint_step(sors_ctx* ctx, const double* q_prev, const double* v_prev, const double* q_next, double dt,
                      const double* qbar_next, const double* vbar_next, double* qbar_prev, double* vbar_prev,
                      double* youngs_bar, double* act_bar, double rtol, int max_iter, long long* pcg_iters);

/* ---- Serving loop --------------------------------------------------------
 * No counterpart in the reference: its analogue is a sequential Python loop
 * over full rollouts (python/d05a_hopping_grid.py:103-106).  A server keeps
 * many rollouts of uneven mesh size resident on the device, admits queued
 * requests in arrival order whenever a slot and HBM budget are free, and
 * advances every resident rollout in lock-step ticks: per-rollout assembly
 * and line-search kernels run on the rollout's own stream while the PCG
 * iterations of all rollouts in their linear solve run as one batched
 * launch per kernel.  Each rollout follows exactly the Newton logic of
 * sors_step with sors_step's predictor (options.predict) and its FP64
 * continuation of a mixed solve that reaches max_newton, so on the sliced
 * node-block layout its result is that of sors_step alone bit for bit. */
typedef struct sors_server sors_server;

typedef struct {
    double hbm_budget_bytes;   /* device memory resident rollouts may hold */
    int max_active;            /* rollouts resident at once (<= 256) */
    int prep_threads;          /* host threads building patterns / uploading */
    int pcg_chunk;             /* batched PCG iterations between host checks (even) */
    sors_options solver;       /* options of every rollout */
} sors_server_options;

typedef struct {
    int status;                /* SORS_OK or the rollout's error code */
    int steps;                 /* implicit steps completed */
    long long newton_iters, pcg_iters, backtracks;
    double latency_ms;         /* submit -> result ready */
    double queue_ms;           /* submit -> admitted */
    double prep_ms;            /* admitted -> resident (pattern, upload) */
    double device_bytes;       /* device memory held while resident */
    int fp64_fallbacks;        /* mixed solves continued in FP64 (as sors_step) */
} sors_request_stats;

typedef struct {
    long long submitted, completed, failed;
    int active, peak_active;
    double resident_bytes, peak_resident_bytes;
    long long ticks, batched_launches;
} sors_server_stats;

void sors_server_default_options(sors_server_options* o);
int sors_server_create(const sors_server_options* o, sors_server** out);
/* Submit one rollout: the mesh and material of desc (copied), initial
 * velocity v0 [3 n_nodes] or NULL, `steps` implicit steps of size dt.
 * Arrival time is the time of the call. */
int sors_server_submit(sors_server* s, const sors_system_desc* desc, const double* v0, double dt, int steps,
                       long long* id);
/* Wait up to timeout_s (< 0: forever) for request `id`; on completion copy
 * its final positions q [3 n_nodes] (may be NULL) and stats, and release
 * it.  Returns the request's status, or SORS_ERR_TIMEOUT. */
int sors_server_wait(sors_server* s, long long id, double* q, sors_request_stats* st, double timeout_s);
int sors_server_get_stats(sors_server* s, sors_server_stats* st);
const char* sors_server_last_error(const sors_server* s);
/* Finishes resident work, drops queued requests, frees everything. */
void sors_server_destroy(sors_server* s);

#ifdef __cplusplus
}
#endif
#endif
