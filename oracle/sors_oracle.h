/*
 * sors_oracle.h -- CPU FP64 restatement of the SORS implicit Newton step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path in
 * sors_amd/: only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it.  Nothing in the product links or calls it.
 *
 * Every function restates reference code (paths relative to the reference
 * repository root, srl-ethz/sors):
 *   cpp/src/geometries/tetrahedron.cpp, cpp/src/geometries/hexahedron.cpp,
 *   cpp/headers/geometries/element.h, cpp/src/energies/neoHookean.cpp,
 *   cpp/src/energies/stableNeoHookean.cpp, cpp/src/energies/softconMuscle.cpp,
 *   cpp/src/energy.cpp, cpp/src/integrator.cpp, cpp/headers/integrator.h,
 *   cpp/src/solver.cpp, cpp/headers/common.h.
 * Citations are given per function in sors_oracle.c.
 *
 * Conventions (same as the reference):
 *   - positions q are FP64 node-major [x0,y0,z0,x1,...] (energy.cpp:216-221);
 *   - element-local DOF = 3*localNode + comp, global DOF = 3*node + comp
 *     (energy.cpp:270-277);
 *   - 3x3 matrices are row-major; "vec" flattens row-major
 *     (neoHookean.cpp:63-64 reshape of the transpose).
 *
 * Linear algebra that the reference takes from Eigen (3x3 inverse, sparse
 * triplet assembly, SimplicialLDLT) is restated here from the published
 * algorithms: cofactor inverse, sort-and-sum triplet assembly, envelope
 * (skyline) LDL^T without pivoting.  Iterate-by-iterate rounding of Eigen's
 * factorisation is not reproduced (only the solution to FP64 accuracy).
 */
#ifndef SORS_ORACLE_H
#define SORS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_BE = 0, ORC_CN = 1 };
/* solverMethod: minimize_newton, minimize_newton_linesearch (with the
 * minimize_newton Dirichlet handling), the SQP assembly semantics (no
 * masking of element terms, no Dirichlet override), and
 * minimize_frozenhessian (the minimize_newton assembly semantics). */
enum { ORC_NEWTON = 0, ORC_NEWTON_LS = 1, ORC_SQP = 2, ORC_FROZEN = 3 };
enum { ORC_LIN_LDLT = 0, ORC_LIN_PCG = 1 };
/* elastic ElementEnergy: "neohookean" (neoHookean.cpp) or "stableneohookean"
 * (stableNeoHookean.cpp) */
enum { ORC_MAT_NH = 0, ORC_MAT_SNH = 1 };
enum {
    ORC_OK = 0,
    ORC_ERR_INVERTED = -1,   /* "Inverted Elements" (neoHookean.cpp:97-101) */
    ORC_ERR_NOCONV = -2,     /* "Newton's method did not converge" (solver.cpp:100-103) */
    ORC_ERR_NAN = -3,        /* dx has NaN (solver.cpp:90-93) */
    ORC_ERR_LS = -4,         /* backtracking cap reached (an addition) */
    ORC_ERR_ARG = -5,
    ORC_ERR_SOLVE = -6       /* zero pivot / PCG breakdown */
};

/* ---- element level ---------------------------------------------------- */
void   orc_lame(double E, double nu, double* mu, double* lam);
int    orc_tet_precompute(const double Xe[12], double Dminv[9], double* V0, double B[108]);
int    orc_hex_precompute(const double Xe[24], double G[192], double B[1728], double* V0);
double orc_nh_psi(const double F[9], double mu, double lam);
void   orc_nh_dpsi(const double F[9], double mu, double lam, double P[9]);
int    orc_nh_d2psi(const double F[9], double mu, double lam, double K[81]);
void   orc_project_psd(int n, double* K);
/* Per-element energy (x V0), gradient (3*ed) and Hessian ((3*ed)^2, row-major).
 * grad / hess may be NULL.  project != 0 clamps each Gauss point's 9x9
 * d2Psi/dF2 to PSD before contraction. */
int    orc_elem_eval(int ed, const double* Xe, const double* xe, double E, double nu,
                     int project, double* energy, double* grad, double* hess);
/* StableNeoHookeanEnergy (stableNeoHookean.cpp:3-150); mu, lam: Lame
 * constants of (E, nu), reparameterised as the reference does */
void   orc_snh_params(double mu, double lam, double* mus, double* lams, double* alpha);
double orc_snh_psi(const double F[9], double mu, double lam);
void   orc_snh_dpsi(const double F[9], double mu, double lam, double P[9]);
void   orc_snh_d2psi(const double F[9], double mu, double lam, double K[81]);
/* orc_elem_eval with an elastic material (ORC_MAT_*) and, when mdir != NULL,
 * the SoftconMuscleEnergy (softconMuscle.cpp:12-134) with direction mdir,
 * stiffness mk and activation mact */
int    orc_elem_eval_ex(int ed, const double* Xe, const double* xe, double E, double nu, int material,
                        const double* mdir, double mk, double mact, int project, double* energy, double* grad,
                        double* hess);

/* ---- system level ----------------------------------------------------- */
typedef struct orc_sys orc_sys;

typedef struct {
    int    linear_solver;   /* ORC_LIN_LDLT | ORC_LIN_PCG */
    int    project;         /* SPD-project element blocks (needed for PCG) */
    int    max_iter;        /* Newton iterations (MAX_ITER = 100, common.h:21) */
    double dx_tol;          /* DX_TOL = 1e-9 (common.h:25) */
    double g_tol;           /* optional residual criterion (<=0: off) */
    double pcg_rtol;        /* relative residual for PCG */
    int    pcg_max_iter;
    int    max_backtrack;   /* cap on Armijo halvings */
    int    nthreads;
} orc_opts;

typedef struct {
    int    newton_iters;
    long long pcg_iters;
    int    backtracks;
    double gnorm;           /* ||g|| at the last evaluated iterate */
    double dxnorm;          /* last accepted step norm */
    double t_grad, t_hess, t_solve, t_energy;   /* seconds */
} orc_stats;

void     orc_default_opts(orc_opts* o);
orc_sys* orc_create(int N, int nel, int ed, const double* X, const int* conn,
                    const double* youngs, const double* poisson, const double* density,
                    const int* bcmask, const double* bcvalue,
                    const double grav[3], double alpha, int scheme, int method);
void     orc_free(orc_sys* s);
int      orc_ndof(const orc_sys* s);
void     orc_get_mass(const orc_sys* s, double* m);
void     orc_set_initial_velocity(orc_sys* s, const double* v);
void     orc_set_initial_deformation(orc_sys* s, const double* q);
void     orc_get_state(const orc_sys* s, double* qprev, double* vprev);
void     orc_set_threads(orc_sys* s, int nthreads);
void     orc_set_external_force(orc_sys* s, const double* f);

double   orc_energy(orc_sys* s, const double* q, double dt);
int      orc_gradient(orc_sys* s, const double* q, double dt, double* g);
/* CSR pattern of the assembled Hessian for this dt (dt==0 vs dt!=0 differ). */
long long orc_pattern(orc_sys* s, double dt);
void     orc_get_pattern(const orc_sys* s, int* rowptr, int* cols);
/* scatter map: nel*(3ed)^2 CSR value indices, -1 for masked pairs */
void     orc_get_scatter_map(const orc_sys* s, int* map);
int      orc_hessian(orc_sys* s, const double* q, double dt, int project, double* vals);
/* Solver::step for one substep-loop call (solver.cpp:5-44, no CFL). */
int      orc_step(orc_sys* s, double* q, double dt, int substeps, const orc_opts* o, orc_stats* st);
void     orc_update_state(orc_sys* s, const double* q, double dt);
/* one minimisation without the state update (minimize_newton[_linesearch],
 * minimize_frozenhessian, minimize_SQP) */
int      orc_minimize(orc_sys* s, double* x, double dt, const orc_opts* o, orc_stats* st);
/* CPU-baseline sample: seconds per FP64 PCG iteration on H(q) (includes the
 * block-Jacobi apply and both dot products; assembly not timed). */
double   orc_bench_pcg(orc_sys* s, const double* q, double dt, int iters, int project);

/* ---- contact (penalty + lagged Coulomb friction) ------------------------
 * A formulation of this build, not of the reference: the reference solves
 * contact as hard constraints n.(x - p) >= 0 by SQP/OSQP.  Only the
 * geometry is the reference's: plane distance n.(x - p)
 * (planeContact.cpp:4-42, plane3D.h:30-33) and the disk's signed plane
 * distance n.(x - c) and radial distance |(x - c) - (n.(x - c)) n|
 * (disk3D.h:33-45, diskContact.cpp:4-42).  See sors_oracle.c for the
 * energy.  surf == NULL: every node is a contact candidate.  Computes the
 * friction lag from the current qprev. */
int  orc_set_contact(orc_sys* s, int nplanes, const double* plane_point, const double* plane_normal, int ndisks,
                     const double* disk_center, const double* disk_normal, const double* disk_radius, int nsurf,
                     const int* surf, double stiffness, double friction, double eps_v);
/* DiskContact::set_disk (diskContact.h:72-80): move disk idx (takes effect at
 * the next step; the displacement since the last step drives friction). */
int  orc_set_disk(orc_sys* s, int idx, const double c[3], const double n[3], double r);
/* contact-only energy / gradient (added into g) at q for the current lag */
double orc_contact_energy(orc_sys* s, const double* q, double dt);
/* recompute the friction lag from qprev (done by orc_step before each step) */
void orc_contact_lag(orc_sys* s, double dt);

/* ---- active fibre stress (tet4): Psi_act = 1/2 sigma a (|F f|^2 - 1) -------
 * A formulation of this build (the reference's actuated element energies
 * take per-group activations the same way: softconMuscle.cpp:12-134). */
int  orc_set_active(orc_sys* s, int ngroups, const int* group, const double* fibre, double sigma);
void orc_set_activation(orc_sys* s, const double* act);
void orc_set_youngs(orc_sys* s, const double* youngs);

/* ---- element energy plugins ------------------------------------------------
 * orc_set_material: the elastic energy of every element (ORC_MAT_*).
 * orc_set_muscle: SoftconMuscleEnergy on the elements with group[e] >= 0,
 * stiffness[e] ("softconStiffness"), dir[3e..3e+2] ("softconDirection"),
 * activations per group (the "softconMuscle" actuation, element.h:52). */
int  orc_set_material(orc_sys* s, int material);
int  orc_set_muscle(orc_sys* s, int ngroups, const int* group, const double* stiffness, const double* dir);
void orc_set_muscle_activation(orc_sys* s, const double* act);

/* ---- PressureForce (cpp/src/forces/pressureForce.cpp:14-185) ---------------
 * Faces of nv = 3 (tet4 meshes) or 4 (hex8) vertex indices, faces[nv f + j],
 * each in an actuation group; pressures per group (the "pressure"
 * actuation).  The force enters the energy as +f(q).(q - qPrev) and the
 * gradient as -f(q) (energy.cpp:228-244, 282-293); its Jacobian enters the
 * Hessian only for minimize_newton (energy.cpp:354-372), with the reference's
 * p/6 factor on both face types.  nfaces = 0 removes the force.
 * orc_pressure_force / orc_pressure_jacobian expose the two pieces: f [n],
 * and the dense Jacobian block list J[f][j][k][3][3] of the reference's
 * triplets (row 3 faces[j] + l, column 3 faces[k] + m). */
int  orc_set_pressure(orc_sys* s, int nfaces, int nv, const int* faces, const int* groups, int ngroups);
void orc_set_pressure_activation(orc_sys* s, const double* p);
void orc_pressure_force(const orc_sys* s, const double* q, double* f);
void orc_pressure_jacobian(const orc_sys* s, const double* q, double* J);

/* ---- minimize_SQP (solver.cpp:109-197) -------------------------------------
 * Under ORC_SQP the planes and disks of orc_set_contact are the reference's
 * planeContact / diskContact inequality constraints and the Dirichlet mask
 * its neumannBC equality constraints; no penalty or friction terms.
 *
 * orc_qp_solve: min 1/2 x^T P x + q^T x  s.t.  l <= A x <= u, P on a CSR
 * pattern (n x n, full symmetric storage, diagonal present), row i of A
 * acting on the DOFs 3 vtx[i] + (0, 1, 2) with coefficients a[3i..3i+2]
 * (u[i] = INFINITY for one-sided rows).  OSQP's ADMM with adaptive rho and
 * solution polishing (the QP solver the reference calls), run until the
 * polished solution is primal feasible with multipliers y of the right sign
 * (OSQP's convention: P x + q + A^T y = 0, y <= 0 on an active lower bound).
 * iters: ADMM iterations.
 *
 * orc_sqp_active: PlaneContact / DiskContact update_active_constraints at x
 * (planeContact.cpp:87-124, diskContact.cpp:91-202): per surface vertex the
 * plane index or -1, the captured disk or -1, and the disk footprint's
 * frozen flag (which persists between calls, as in the reference). */
int  orc_qp_solve(int n, const int* rowptr, const int* cols, const double* P, const double* q, int m, const int* vtx,
                  const double* a, const double* l, const double* u, double* x, double* y, int* iters);
int  orc_sqp_active(orc_sys* s, const double* x, int* pmask, int* dmask, int* frozen);

#ifdef __cplusplus
}
#endif
#endif
