/*
 * sors_oracle.c -- CPU FP64 restatement of the SORS implicit Newton step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the thing measured or
 * shipped).  See sors_oracle.h for the conventions and the list of
 * reference files restated here.
 *
 * Pinning (what checks each function; tests/test_oracle_*.py):
 *   orc_tet_precompute      cpp/src/_test.cpp:217-235 known-answer dF/dx;
 *                           closed-form volume; Dm * Dm^-1 = I
 *   orc_hex_precompute      axis-aligned box: analytic shape gradients,
 *                           F(x = X) = I at all Gauss points, V0 = box volume
 *   orc_nh_psi/dpsi/d2psi   cpp/src/_test.cpp:238-286 (zero at rest);
 *                           torch FP64 autograd golden vectors; central FD
 *   orc_project_psd         numpy.linalg.eigh clamp
 *   orc_elem_eval           torch autograd of Psi(F(x)) per element
 *   orc_gradient/energy     cpp/src/_test.cpp:136-159 (Testahedron energy =
 *                           3, gradient ground truth, via the polynomial
 *                           test element), FD of the energy
 *   orc_pattern/hessian     brute-force triplet sort+sum on small meshes;
 *                           _test.cpp:162-176 diagonal ground truth
 *   orc_step                _test.cpp:395-443 (free fall -10 m/s^2 +-1e-3),
 *                           _test.cpp:180-199 (||g||^2 < 1e-9); LDL^T vs PCG
 *                           agree on converged states
 */
#include "sors_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_TEST_POLY (-4) /* ed value selecting the polynomial test element */

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ------------------------------------------------------------------------ */
/* 3x3 helpers (row-major)                                                   */
/* ------------------------------------------------------------------------ */
static double det3(const double* A) {
    return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
           A[2] * (A[3] * A[7] - A[4] * A[6]);
}

/* inverse by cofactors (Eigen's fixed-size 3x3 inverse is cofactor-based) */
static double inv3(const double* A, double* Ai) {
    double d = det3(A);
    double id = 1.0 / d;
    Ai[0] = (A[4] * A[8] - A[5] * A[7]) * id;
    Ai[1] = (A[2] * A[7] - A[1] * A[8]) * id;
    Ai[2] = (A[1] * A[5] - A[2] * A[4]) * id;
    Ai[3] = (A[5] * A[6] - A[3] * A[8]) * id;
    Ai[4] = (A[0] * A[8] - A[2] * A[6]) * id;
    Ai[5] = (A[2] * A[3] - A[0] * A[5]) * id;
    Ai[6] = (A[3] * A[7] - A[4] * A[6]) * id;
    Ai[7] = (A[1] * A[6] - A[0] * A[7]) * id;
    Ai[8] = (A[0] * A[4] - A[1] * A[3]) * id;
    return d;
}

static void transpose3(const double* A, double* At) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) At[3 * j + i] = A[3 * i + j];
}

/* ------------------------------------------------------------------------ */
/* Element level                                                             */
/* ------------------------------------------------------------------------ */

/* Lame constants: neoHookean.cpp:10-11 */
void orc_lame(double E, double nu, double* mu, double* lam) {
    *lam = (E * nu) / ((1 + nu) * (1 - 2 * nu));
    *mu = E / (2 * (1 + nu));
}

/* volume of a tet: tetrahedron.cpp:137-150 (|(b-a).((c-a)x(d-a))|/6) */
static double tet_volume(const double* a, const double* b, const double* c, const double* d) {
    double ab[3], ac[3], ad[3];
    for (int k = 0; k < 3; ++k) {
        ab[k] = b[k] - a[k];
        ac[k] = c[k] - a[k];
        ad[k] = d[k] - a[k];
    }
    double cx = ac[1] * ad[2] - ac[2] * ad[1];
    double cy = ac[2] * ad[0] - ac[0] * ad[2];
    double cz = ac[0] * ad[1] - ac[1] * ad[0];
    return fabs(ab[0] * cx + ab[1] * cy + ab[2] * cz) / 6.0;
}

/* tet 9x12 dF/dx from Dm^-1: tetrahedron.cpp:17-37.  Column i<9 is
 * vec(E_{i%3, i/3} Dm^-1); columns 9..11 are vec(R_m Dm^-1) with R_m a row
 * of -1 at row m.  Row-major 9x12. */
static void tet_dfdx(const double* Dminv, double* B) {
    memset(B, 0, sizeof(double) * 108);
    for (int i = 0; i < 9; ++i) {
        int r = i % 3, node = i / 3;
        for (int c = 0; c < 3; ++c) B[(3 * r + c) * 12 + i] = Dminv[3 * node + c];
    }
    for (int m = 0; m < 3; ++m)
        for (int c = 0; c < 3; ++c)
            B[(3 * m + c) * 12 + 9 + m] = -(Dminv[c] + Dminv[3 + c] + Dminv[6 + c]);
}

/* Tetrahedron constructor: tetrahedron.cpp:5-37 + compute_volume :137-150 */
int orc_tet_precompute(const double Xe[12], double Dminv[9], double* V0, double B[108]) {
    double Dm[9];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) Dm[3 * r + c] = Xe[3 * c + r] - Xe[9 + r];
    double d = inv3(Dm, Dminv);
    *V0 = tet_volume(Xe, Xe + 3, Xe + 6, Xe + 9);
    if (B) tet_dfdx(Dminv, B);
    return (d != 0.0 && *V0 > 0.0) ? ORC_OK : ORC_ERR_ARG;
}

/* reference cube nodes: hexahedron.cpp:171-179 */
static const double HEX_XI[8][3] = {{1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, -1},
                                    {1, -1, 1},  {1, 1, 1},  {-1, 1, 1},  {-1, -1, 1}};
/* 5-tet split used for the hex volume: hexahedron.cpp:106-130 */
static const int HEX_TETS[5][4] = {{0, 1, 3, 4}, {1, 2, 3, 6}, {1, 4, 5, 6}, {3, 4, 6, 7}, {1, 3, 4, 6}};

/* Hexahedron constructor: compute_shape_functions_gradients (hexahedron.cpp:170-207),
 * compute_deformation_hessians (:145-168), compute_volume (:106-130).
 * G is [s][a][c] (8 x 8 x 3); B is [s] 9x24 row-major. */
int orc_hex_precompute(const double Xe[24], double G[192], double B[1728], double* V0) {
    const double g = 1.0 / sqrt(3.0);
    int ok = ORC_OK;
    for (int s = 0; s < 8; ++s) {
        double qs[3] = {g * HEX_XI[s][0], g * HEX_XI[s][1], g * HEX_XI[s][2]};
        double hatG[8][3];
        for (int i = 0; i < 8; ++i) {
            const double* ni = HEX_XI[i];
            hatG[i][0] = 0.125 * (ni[0] * (1 + ni[1] * qs[1]) * (1 + ni[2] * qs[2]));
            hatG[i][1] = 0.125 * (ni[1] * (1 + ni[0] * qs[0]) * (1 + ni[2] * qs[2]));
            hatG[i][2] = 0.125 * (ni[2] * (1 + ni[0] * qs[0]) * (1 + ni[1] * qs[1]));
        }
        /* hatF = X^T hatG (3x3) */
        double hatF[9] = {0}, hatFi[9];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double acc = 0;
                for (int i = 0; i < 8; ++i) acc += Xe[3 * i + r] * hatG[i][c];
                hatF[3 * r + c] = acc;
            }
        if (inv3(hatF, hatFi) == 0.0) ok = ORC_ERR_ARG;
        /* G_s = (hatF^-T hatG^T)^T = hatG hatF^-1 */
        for (int a = 0; a < 8; ++a)
            for (int c = 0; c < 3; ++c) {
                double acc = 0;
                for (int k = 0; k < 3; ++k) acc += hatG[a][k] * hatFi[3 * k + c];
                G[s * 24 + 3 * a + c] = acc;
            }
        if (B) {
            double* Bs = B + s * 216;
            memset(Bs, 0, sizeof(double) * 216);
            /* column i = 3a + j: vec(E_{j,a} G_s) (3x8 unit times 8x3) */
            for (int a = 0; a < 8; ++a)
                for (int j = 0; j < 3; ++j)
                    for (int c = 0; c < 3; ++c) Bs[(3 * j + c) * 24 + 3 * a + j] = G[s * 24 + 3 * a + c];
        }
    }
    double vol = 0;
    for (int t = 0; t < 5; ++t) {
        double v = tet_volume(Xe + 3 * HEX_TETS[t][0], Xe + 3 * HEX_TETS[t][1], Xe + 3 * HEX_TETS[t][2],
                              Xe + 3 * HEX_TETS[t][3]);
        if (!(v > 0)) ok = ORC_ERR_ARG;
        vol += v;
    }
    *V0 = vol;
    return ok;
}

/* NeoHookeanEnergy::compute_energy per quadrature point: neoHookean.cpp:31-37 */
double orc_nh_psi(const double F[9], double mu, double lam) {
    double J = det3(F);
    double logJ = log(J);
    double trFtF = 0;
    for (int i = 0; i < 9; ++i) trFtF += F[i] * F[i];
    return 0.5 * mu * (trFtF - 3) - mu * logJ + 0.5 * lam * logJ * logJ;
}

/* dE/dF (row-major vec): neoHookean.cpp:61-68 */
void orc_nh_dpsi(const double F[9], double mu, double lam, double P[9]) {
    double Fi[9], FiT[9];
    double J = inv3(F, Fi);
    transpose3(Fi, FiT);
    double logJ = log(J);
    for (int i = 0; i < 9; ++i) P[i] = mu * F[i] + (lam * logJ - mu) * FiT[i];
}

/* d2E/dF2 (9x9): neoHookean.cpp:93-125.
 *   K = mu I + lam f f^T + (lam logJ - mu) FinvTGrad,
 *   FinvTGrad(:, 3p+q) = vec(-F^-T E_pq^T F^-T)  (unitMatrices_ (p,q), :46-53). */
int orc_nh_d2psi(const double F[9], double mu, double lam, double K[81]) {
    double Fi[9], FiT[9];
    double J = inv3(F, Fi);
    transpose3(Fi, FiT);
    if (J <= 0) return ORC_ERR_INVERTED;
    double logJ = log(J);
    double c = lam * logJ - mu;
    for (int r = 0; r < 9; ++r)
        for (int col = 0; col < 9; ++col) {
            int a = r / 3, b = r % 3, p = col / 3, q = col % 3;
            double gradFinvT = -FiT[3 * a + q] * FiT[3 * p + b];
            K[9 * r + col] = (r == col ? mu : 0.0) + lam * FiT[r] * FiT[col] + c * gradFinvT;
        }
    return ORC_OK;
}

/* StableNeoHookeanEnergy constructor (stableNeoHookean.cpp:9-14) from the
 * Lame constants of (E, nu): mu_ = 4/3 mu, lambda_ = lam + 5/8 mu_,
 * alpha_ = 1 + 3/4 mu_ / lambda_. */
void orc_snh_params(double mu, double lam, double* mus, double* lams, double* alpha) {
    *mus = (4.0 / 3.0) * mu;
    *lams = lam + (5.0 / 8.0) * (*mus);
    *alpha = 1 + (3.0 / 4.0) * ((*mus) / (*lams));
}

/* StableNeoHookeanEnergy::compute_energy per quadrature point
 * (stableNeoHookean.cpp:31-40) */
double orc_snh_psi(const double F[9], double mu, double lam) {
    double m, l, al;
    orc_snh_params(mu, lam, &m, &l, &al);
    double J = det3(F);
    double IC = 0;
    for (int i = 0; i < 9; ++i) IC += F[i] * F[i];
    return (m / 2.0) * (IC - 3.0) + (l / 2.0) * (J - al) * (J - al) - (m / 2.0) * log(IC + 1.0);
}

/* compute_gradient per quadrature point (stableNeoHookean.cpp:64-77) */
void orc_snh_dpsi(const double F[9], double mu, double lam, double P[9]) {
    double m, l, al, Fi[9], FiT[9];
    orc_snh_params(mu, lam, &m, &l, &al);
    double J = inv3(F, Fi);
    transpose3(Fi, FiT);
    double IC = 0;
    for (int i = 0; i < 9; ++i) IC += F[i] * F[i];
    for (int i = 0; i < 9; ++i) P[i] = m * F[i] + l * (J - al) * J * FiT[i] - (m * F[i]) / (IC + 1.0);
}

/* compute_hessian per quadrature point (stableNeoHookean.cpp:100-131):
 * A = mu_ I + (2J - alpha) lambda_ J f f^T + lambda_ (J - alpha) J FinvTGrad
 *     - mu_ I / (IC + 1) + 2 mu_ F F^T / (IC + 1)^2 */
void orc_snh_d2psi(const double F[9], double mu, double lam, double K[81]) {
    double m, l, al, Fi[9], FiT[9];
    orc_snh_params(mu, lam, &m, &l, &al);
    double J = inv3(F, Fi);
    transpose3(Fi, FiT);
    double IC = 0;
    for (int i = 0; i < 9; ++i) IC += F[i] * F[i];
    for (int r = 0; r < 9; ++r)
        for (int col = 0; col < 9; ++col) {
            int a = r / 3, b = r % 3, p = col / 3, q = col % 3;
            double gradFinvT = -FiT[3 * a + q] * FiT[3 * p + b];
            double id = (r == col) ? 1.0 : 0.0;
            K[9 * r + col] = m * id + (2 * J - al) * l * J * FiT[r] * FiT[col] + l * (J - al) * J * gradFinvT +
                             ((-m * id) / (IC + 1.0)) + ((2 * m * F[r] * F[col]) / ((IC + 1.0) * (IC + 1.0)));
        }
}

/* Cyclic Jacobi eigen-decomposition of a symmetric n x n matrix, negative
 * eigenvalues clamped to zero, matrix rebuilt in place (textbook method,
 * Golub & Van Loan 8.5).  Used to SPD-project the 9x9 d2Psi/dF2. */
void orc_project_psd(int n, double* K) {
    double A[81 * 9], V[81 * 9];
    if (n > 27) return;
    memcpy(A, K, sizeof(double) * n * n);
    for (int i = 0; i < n * n; ++i) V[i] = 0;
    for (int i = 0; i < n; ++i) V[i * n + i] = 1;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0, tot = 0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                tot += A[i * n + j] * A[i * n + j];
                if (i != j) off += A[i * n + j] * A[i * n + j];
            }
        if (off <= 1e-30 * tot || off == 0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = A[p * n + q];
                if (apq == 0) continue;
                double app = A[p * n + p], aqq = A[q * n + q];
                double theta = (aqq - app) / (2 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
                double cs = 1 / sqrt(t * t + 1), sn = t * cs;
                for (int k = 0; k < n; ++k) {
                    double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = cs * akp - sn * akq;
                    A[k * n + q] = sn * akp + cs * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = cs * apk - sn * aqk;
                    A[q * n + k] = sn * apk + cs * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = cs * vkp - sn * vkq;
                    V[k * n + q] = sn * vkp + cs * vkq;
                }
            }
    }
    for (int i = 0; i < n * n; ++i) K[i] = 0;
    for (int k = 0; k < n; ++k) {
        double lk = A[k * n + k];
        if (lk <= 0) continue;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) K[i * n + j] += lk * V[i * n + k] * V[j * n + k];
    }
}

/* Polynomial test element (the reference's Testahedron, cpp/src/_test.cpp:8-66):
 * psi = sum x^2 (x^2 - 0.5) over the element's 12 coordinates. */
static void poly_eval(const double* xe, double* energy, double* grad, double* hess) {
    if (energy) {
        double e = 0;
        for (int i = 0; i < 12; ++i) e += xe[i] * xe[i] * (xe[i] * xe[i] - 0.5);
        *energy = e;
    }
    if (grad)
        for (int i = 0; i < 12; ++i) grad[i] = (2 * xe[i] - 1) * (2 * xe[i] + 1) * xe[i];
    if (hess) {
        memset(hess, 0, sizeof(double) * 144);
        for (int i = 0; i < 12; ++i) hess[13 * i] = 12 * xe[i] * xe[i] - 1;
    }
}

/* Element::compute_energy (element.h:40-58) / Tetrahedron::compute_gradient,
 * compute_hessian (tetrahedron.cpp:99-134) / Hexahedron (hexahedron.cpp:68-103)
 * with the NeoHookeanEnergy quadrature (weights 1/nquad, neoHookean.cpp:26-27,
 * 56-57, 85-86).  Literal dense form: grad = V0 sum_s w dEdF_s B_s,
 * hess = V0 sum_s w B_s^T K_s B_s. */
/* sa, f: active fibre stress sigma_max * a and unit fibre f (f == NULL or
 * sa == 0: passive).  Psi_act = 1/2 sa (|F f|^2 - 1), P_act = sa (F f) f^T,
 * dP_act/dF = sa I (x) f f^T (PSD, added after any projection). */
/* SoftconMuscleEnergy at one quadrature point (softconMuscle.cpp:12-134):
 * F, B = nablaF (9 x nd), direction m, stiffness k, activation e.  Adds
 * w * (energy, gradient, Hessian).  project != 0: the Hessian is
 * B^T K B with K = k [(r/l^2) vec(u m^T) vec(u m^T)^T + (1-r) vec(1 m^T)
 * vec(1 m^T)^T] (the reference's Hessian in 9x9 form) clamped to PSD. */
static void muscle_point(int nd, const double* F, const double* B, const double* m, double k, double e, double w,
                         int project, double* energy, double* grad, double* hess) {
    double Fm[3];
    for (int a = 0; a < 3; ++a) Fm[a] = F[3 * a] * m[0] + F[3 * a + 1] * m[1] + F[3 * a + 2] * m[2];
    double l = sqrt(Fm[0] * Fm[0] + Fm[1] * Fm[1] + Fm[2] * Fm[2]);
    double r = (1 - e) / l;
    if (energy) {
        /* k_ / 2 * ((1 - r) * F * m_).norm()  (:30-31) */
        double v0 = (1 - r) * Fm[0], v1 = (1 - r) * Fm[1], v2 = (1 - r) * Fm[2];
        *energy += (k / 2 * sqrt(v0 * v0 + v1 * v1 + v2 * v2)) * w;
    }
    if (!grad && !hess) return;
    double t1[24], t2[24];
    for (int i = 0; i < nd; ++i) {
        /* m_c F_ab m_b nablaF_(ac),i  (:64-72) and nablaF_(ac),i m_c (:108-115) */
        double a1 = 0, a2 = 0;
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) {
                a1 += m[c] * Fm[a] * B[(3 * a + c) * nd + i];
                a2 += B[(3 * a + c) * nd + i] * m[c];
            }
        t1[i] = a1;
        t2[i] = a2;
    }
    if (grad)
        for (int i = 0; i < nd; ++i) grad[i] += t1[i] * k * (1 - r) * w;
    if (!hess) return;
    if (!project) {
        for (int i = 0; i < nd; ++i)
            for (int j = 0; j < nd; ++j)
                hess[i * nd + j] += k * ((r / (l * l)) * t1[i] * t1[j] + (1 - r) * t2[i] * t2[j]) * w;
        return;
    }
    double K[81], v1[9], v2[9];
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) {
            v1[3 * a + c] = Fm[a] * m[c];
            v2[3 * a + c] = m[c];
        }
    for (int p = 0; p < 9; ++p)
        for (int q = 0; q < 9; ++q) K[9 * p + q] = k * ((r / (l * l)) * v1[p] * v1[q] + (1 - r) * v2[p] * v2[q]);
    orc_project_psd(9, K);
    for (int i = 0; i < nd; ++i)
        for (int j = 0; j < nd; ++j) {
            double acc = 0;
            for (int p = 0; p < 9; ++p)
                for (int q = 0; q < 9; ++q) acc += B[p * nd + i] * K[9 * p + q] * B[q * nd + j];
            hess[i * nd + j] += acc * w;
        }
}

/* material: ORC_MAT_NH (neoHookean.cpp) or ORC_MAT_SNH (stableNeoHookean.cpp);
 * mdir != NULL adds the SoftconMuscle energy with stiffness mk and
 * activation mact.  Each energy's Hessian is projected on its own. */
static int elem_eval_pre(int ed, const double* pre, double V0, const double* xe, double mu, double lam,
                         int project, double* energy, double* grad, double* hess, double sa, const double* f,
                         int material, const double* mdir, double mk, double mact) {
    const int active = f != NULL && sa != 0.0;
    const int snh = material == ORC_MAT_SNH;
    int nq = (ed == 4) ? 1 : 8;
    int nd = 3 * ed;
    double w = 1.0 / nq;
    double Bbuf[216];
    double e = 0;
    if (grad) memset(grad, 0, sizeof(double) * nd);
    if (hess) memset(hess, 0, sizeof(double) * nd * nd);
    int status = ORC_OK;
    for (int s = 0; s < nq; ++s) {
        double F[9];
        const double* Gs; /* per-node gradient vectors: node a -> Gs[3a..3a+2] */
        double gtet[12];
        if (ed == 4) {
            /* F = Ds Dm^-1 (tetrahedron.cpp:152-169) */
            const double* Dminv = pre;
            double Ds[9];
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r) Ds[3 * r + c] = xe[3 * c + r] - xe[9 + r];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c)
                    F[3 * r + c] = Ds[3 * r] * Dminv[c] + Ds[3 * r + 1] * Dminv[3 + c] + Ds[3 * r + 2] * Dminv[6 + c];
            tet_dfdx(Dminv, Bbuf);
            (void)gtet;
            Gs = NULL;
        } else {
            /* F_s = x^T G_s (hexahedron.cpp:133-142) */
            Gs = pre + 24 * s;
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) {
                    double acc = 0;
                    for (int a = 0; a < 8; ++a) acc += xe[3 * a + r] * Gs[3 * a + c];
                    F[3 * r + c] = acc;
                }
            memset(Bbuf, 0, sizeof(double) * 216);
            for (int a = 0; a < 8; ++a)
                for (int j = 0; j < 3; ++j)
                    for (int c = 0; c < 3; ++c) Bbuf[(3 * j + c) * 24 + 3 * a + j] = Gs[3 * a + c];
        }
        double Ff[3] = {0, 0, 0};
        if (active)
            for (int r = 0; r < 3; ++r) Ff[r] = F[3 * r] * f[0] + F[3 * r + 1] * f[1] + F[3 * r + 2] * f[2];
        if (energy) {
            e += (snh ? orc_snh_psi(F, mu, lam) : orc_nh_psi(F, mu, lam)) * w;
            if (active) e += 0.5 * sa * (Ff[0] * Ff[0] + Ff[1] * Ff[1] + Ff[2] * Ff[2] - 1.0) * w;
        }
        if (mdir) muscle_point(nd, F, Bbuf, mdir, mk, mact, w, project, energy ? &e : NULL, grad, hess);
        if (grad) {
            double P[9];
            if (snh)
                orc_snh_dpsi(F, mu, lam, P);
            else
                orc_nh_dpsi(F, mu, lam, P);
            if (active)
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) P[3 * r + c] += sa * Ff[r] * f[c];
            for (int i = 0; i < nd; ++i) {
                double acc = 0;
                for (int r = 0; r < 9; ++r) acc += P[r] * Bbuf[r * nd + i];
                grad[i] += acc * w;
            }
        }
        if (hess) {
            double K[81], KB[9 * 24];
            if (snh) {
                orc_snh_d2psi(F, mu, lam, K);
            } else {
                int st = orc_nh_d2psi(F, mu, lam, K);
                if (st != ORC_OK) status = st;
            }
            if (project) orc_project_psd(9, K);
            if (active)
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j)
                        for (int l = 0; l < 3; ++l) K[9 * (3 * i + j) + 3 * i + l] += sa * f[j] * f[l];
            for (int r = 0; r < 9; ++r)
                for (int i = 0; i < nd; ++i) {
                    double acc = 0;
                    for (int k = 0; k < 9; ++k) acc += K[9 * r + k] * Bbuf[k * nd + i];
                    KB[r * nd + i] = acc;
                }
            for (int i = 0; i < nd; ++i)
                for (int j = 0; j < nd; ++j) {
                    double acc = 0;
                    for (int r = 0; r < 9; ++r) acc += Bbuf[r * nd + i] * KB[r * nd + j];
                    hess[i * nd + j] += acc * w;
                }
        }
    }
    if (energy) *energy = e * V0;
    if (grad)
        for (int i = 0; i < nd; ++i) grad[i] *= V0;
    if (hess)
        for (int i = 0; i < nd * nd; ++i) hess[i] *= V0;
    return status;
}

int orc_elem_eval(int ed, const double* Xe, const double* xe, double E, double nu, int project,
                  double* energy, double* grad, double* hess) {
    double mu, lam, V0;
    if (ed == ORC_TEST_POLY) {
        poly_eval(xe, energy, grad, hess);
        return ORC_OK;
    }
    orc_lame(E, nu, &mu, &lam);
    if (ed == 4) {
        double Dminv[9];
        int st = orc_tet_precompute(Xe, Dminv, &V0, NULL);
        if (st) return st;
        return elem_eval_pre(4, Dminv, V0, xe, mu, lam, project, energy, grad, hess, 0.0, NULL, ORC_MAT_NH, NULL,
                             0.0, 0.0);
    }
    if (ed == 8) {
        double G[192];
        int st = orc_hex_precompute(Xe, G, NULL, &V0);
        if (st) return st;
        return elem_eval_pre(8, G, V0, xe, mu, lam, project, energy, grad, hess, 0.0, NULL, ORC_MAT_NH, NULL, 0.0,
                             0.0);
    }
    return ORC_ERR_ARG;
}

int orc_elem_eval_ex(int ed, const double* Xe, const double* xe, double E, double nu, int material,
                     const double* mdir, double mk, double mact, int project, double* energy, double* grad,
                     double* hess) {
    double mu, lam, V0;
    orc_lame(E, nu, &mu, &lam);
    if (ed == 4) {
        double Dminv[9];
        int st = orc_tet_precompute(Xe, Dminv, &V0, NULL);
        if (st) return st;
        return elem_eval_pre(4, Dminv, V0, xe, mu, lam, project, energy, grad, hess, 0.0, NULL, material, mdir, mk,
                             mact);
    }
    if (ed == 8) {
        double G[192];
        int st = orc_hex_precompute(Xe, G, NULL, &V0);
        if (st) return st;
        return elem_eval_pre(8, G, V0, xe, mu, lam, project, energy, grad, hess, 0.0, NULL, material, mdir, mk,
                             mact);
    }
    return ORC_ERR_ARG;
}

/* ------------------------------------------------------------------------ */
/* System level                                                              */
/* ------------------------------------------------------------------------ */
#define ORC_MAX_PRIM 4
#define ORC_DISK_RIM 1e-3   /* rim falloff width: the reference's radial threshold */
struct orc_sys {
    int N, nel, ed, nd, n, method, scheme, poly;
    int* conn;
    double* pre;   /* per element: Dm^-1 (9) or G (192) */
    double* V0;
    double* mu;
    double* lam;
    double* mass;  /* per DOF, energy.cpp:176-189 */
    int* mask;
    double* bcval;
    double grav[3];
    double alpha;
    double *qprev, *vprev, *gprev, *gprevtmp;
    /* pattern */
    double* fext;  /* external vertex force (VertexForce), NULL if none */
    int pattern_dtnz; /* -1: none built */
    int* rowptr;
    int* cols;
    long long nnz;
    int nthreads;
    /* contact (orc_set_contact) */
    int nplanes, ndisks, nsurf, has_contact;
    double plane_p[ORC_MAX_PRIM][3], plane_n[ORC_MAX_PRIM][3];
    double disk_c[ORC_MAX_PRIM][3], disk_n[ORC_MAX_PRIM][3], disk_r[ORC_MAX_PRIM], disk_c_last[ORC_MAX_PRIM][3];
    int* surf;
    double kc, muf, eps_v, lag_dt;
    /* active fibre stress (orc_set_active) */
    int ngroups;
    int* agroup;
    double* afib;   /* [nel][3] */
    double* act;    /* [ngroups] */
    double sigma;
    double* nu;     /* Poisson ratio per element (orc_set_youngs) */
    int material;   /* ORC_MAT_NH | ORC_MAT_SNH */
    /* SoftconMuscle (orc_set_muscle) */
    int mngroups;
    int* mgroup;    /* [nel] group or -1 */
    double* mk;     /* [nel] softconStiffness */
    double* mdir;   /* [nel][3] softconDirection */
    double* mact;   /* [mngroups] activation */
    /* PressureForce (orc_set_pressure) */
    int npf, pnv, pngroups;
    int* pfaces;    /* [npf][pnv] */
    int* pgroup;    /* [npf] */
    double* pact;   /* [pngroups] pressure */
    double* lagN;   /* [nsurf] lagged normal force magnitude */
    double* lagn;   /* [nsurf][3] its normal */
    double* lagD;   /* [nsurf][3] displacement of its primitive over the step */
    /* minimize_SQP active sets: closest plane / captured disk per surface
     * vertex or -1 (planeCollisionMask_, diskCollisionMask_) and the disk
     * footprint's frozen flag (frozenActiveSet_) */
    int* pmask;
    int* dmask;
    int dfrozen;
};

void orc_default_opts(orc_opts* o) {
    o->linear_solver = ORC_LIN_LDLT;
    o->project = 0;
    o->max_iter = 100;
    o->dx_tol = 1e-9;
    o->g_tol = 0;
    o->pcg_rtol = 1e-10;
    o->pcg_max_iter = 20000;
    o->max_backtrack = 60;
    o->nthreads = 1;
}

/* Energy::Energy (energy.cpp:4-204): element precompute, lumped mass
 * (:176-189), Dirichlet mask (:61-68), TimeIntegrator(qPrev = X, vPrev = 0)
 * (integrator.h:155-161). ed == -4 selects the polynomial test element. */
orc_sys* orc_create(int N, int nel, int ed, const double* X, const int* conn, const double* youngs,
                    const double* poisson, const double* density, const int* bcmask, const double* bcvalue,
                    const double grav[3], double alpha, int scheme, int method) {
    int poly = (ed == ORC_TEST_POLY);
    if (poly) ed = 4;
    if (ed != 4 && ed != 8) return NULL;
    orc_sys* s = (orc_sys*)calloc(1, sizeof(orc_sys));
    s->N = N;
    s->nel = nel;
    s->ed = ed;
    s->nd = 3 * ed;
    s->n = 3 * N;
    s->method = method;
    s->scheme = scheme;
    s->poly = poly;
    s->nthreads = 1;
    s->conn = (int*)malloc(sizeof(int) * (size_t)nel * ed);
    memcpy(s->conn, conn, sizeof(int) * (size_t)nel * ed);
    int psz = (ed == 4) ? 9 : 192;
    s->pre = (double*)malloc(sizeof(double) * (size_t)nel * psz);
    s->V0 = (double*)malloc(sizeof(double) * nel);
    s->mu = (double*)malloc(sizeof(double) * nel);
    s->lam = (double*)malloc(sizeof(double) * nel);
    s->nu = (double*)malloc(sizeof(double) * (nel > 0 ? nel : 1));
    s->mass = (double*)calloc(s->n, sizeof(double));
    s->mask = (int*)calloc(s->n, sizeof(int));
    s->bcval = (double*)calloc(s->n, sizeof(double));
    s->qprev = (double*)malloc(sizeof(double) * s->n);
    s->vprev = (double*)calloc(s->n, sizeof(double));
    s->gprev = (double*)calloc(s->n, sizeof(double));
    s->gprevtmp = (double*)calloc(s->n, sizeof(double));
    s->rowptr = NULL;
    s->cols = NULL;
    s->pattern_dtnz = -1;
    for (int i = 0; i < 3; ++i) s->grav[i] = grav ? grav[i] : 0.0;
    s->alpha = alpha;
    if (bcmask)
        for (int i = 0; i < s->n; ++i) s->mask[i] = bcmask[i] != 0;
    if (bcvalue) memcpy(s->bcval, bcvalue, sizeof(double) * s->n);
    memcpy(s->qprev, X, sizeof(double) * s->n);
    for (int e = 0; e < nel; ++e) {
        double Xe[24];
        for (int a = 0; a < ed; ++a)
            for (int k = 0; k < 3; ++k) Xe[3 * a + k] = X[3 * (size_t)conn[(size_t)e * ed + a] + k];
        if (ed == 4)
            orc_tet_precompute(Xe, s->pre + (size_t)e * 9, &s->V0[e], NULL);
        else
            orc_hex_precompute(Xe, s->pre + (size_t)e * 192, NULL, &s->V0[e]);
        orc_lame(youngs ? youngs[e] : 0.0, poisson ? poisson[e] : 0.0, &s->mu[e], &s->lam[e]);
        s->nu[e] = poisson ? poisson[e] : 0.0;
        double me = (density ? density[e] : 0.0) * s->V0[e]; /* element.h:31 */
        for (int a = 0; a < ed; ++a)
            for (int k = 0; k < 3; ++k) s->mass[3 * (size_t)conn[(size_t)e * ed + a] + k] += me / ed;
    }
    return s;
}

void orc_free(orc_sys* s) {
    if (!s) return;
    free(s->conn);
    free(s->pre);
    free(s->V0);
    free(s->mu);
    free(s->lam);
    free(s->mass);
    free(s->mask);
    free(s->bcval);
    free(s->qprev);
    free(s->vprev);
    free(s->gprev);
    free(s->gprevtmp);
    free(s->rowptr);
    free(s->cols);
    free(s->fext);
    free(s->nu);
    free(s->agroup);
    free(s->afib);
    free(s->act);
    free(s->mgroup);
    free(s->mk);
    free(s->mdir);
    free(s->mact);
    free(s->pfaces);
    free(s->pgroup);
    free(s->pact);
    free(s->surf);
    free(s->lagN);
    free(s->lagn);
    free(s->lagD);
    free(s->pmask);
    free(s->dmask);
    free(s);
}

/* VertexForce::compute_force returns the actuation vector (vertexForce.cpp:4-8);
 * Energy adds f.(q - qPrev) to the energy (energy.cpp:228-244) and subtracts f
 * from the gradient (:282-293).  f == NULL removes the force. */
void orc_set_external_force(orc_sys* s, const double* f) {
    if (!f) {
        free(s->fext);
        s->fext = NULL;
        return;
    }
    if (!s->fext) s->fext = (double*)malloc(sizeof(double) * s->n);
    memcpy(s->fext, f, sizeof(double) * s->n);
}

/* PressureForce::PressureForce (pressureForce.cpp:4-11) */
int orc_set_pressure(orc_sys* s, int nfaces, int nv, const int* faces, const int* groups, int ngroups) {
    free(s->pfaces);
    free(s->pgroup);
    free(s->pact);
    s->pfaces = NULL;
    s->pgroup = NULL;
    s->pact = NULL;
    s->npf = 0;
    if (nfaces <= 0) return ORC_OK;
    if (nv != (s->ed == 8 ? 4 : 3) || ngroups < 1) return ORC_ERR_ARG;
    for (int f = 0; f < nfaces; ++f) {
        if (groups[f] < 0 || groups[f] >= ngroups) return ORC_ERR_ARG;
        for (int j = 0; j < nv; ++j)
            if (faces[(size_t)nv * f + j] < 0 || faces[(size_t)nv * f + j] >= s->N) return ORC_ERR_ARG;
    }
    s->npf = nfaces;
    s->pnv = nv;
    s->pngroups = ngroups;
    s->pfaces = (int*)malloc(sizeof(int) * (size_t)nv * nfaces);
    s->pgroup = (int*)malloc(sizeof(int) * (size_t)nfaces);
    s->pact = (double*)calloc((size_t)ngroups, sizeof(double));
    memcpy(s->pfaces, faces, sizeof(int) * (size_t)nv * nfaces);
    memcpy(s->pgroup, groups, sizeof(int) * (size_t)nfaces);
    return ORC_OK;
}

void orc_set_pressure_activation(orc_sys* s, const double* p) {
    if (s->npf > 0) memcpy(s->pact, p, sizeof(double) * s->pngroups);
}

/* area vector of a face at q: a triangle's (x1 - x0) x (x2 - x0) / 2, a
 * quad's two triangles (x0, x1, x3) and (x2, x3, x1) (pressureForce.cpp:22-33) */
static void face_area(const orc_sys* s, int f, const double* q, double A[3]) {
    const int* fv = s->pfaces + (size_t)s->pnv * f;
    const double* x0 = q + 3 * (size_t)fv[0];
    const double* x1 = q + 3 * (size_t)fv[1];
    if (s->pnv == 3) {
        const double* x2 = q + 3 * (size_t)fv[2];
        double a[3], b[3];
        for (int k = 0; k < 3; ++k) {
            a[k] = x1[k] - x0[k];
            b[k] = x2[k] - x0[k];
        }
        A[0] = 0.5 * (a[1] * b[2] - a[2] * b[1]);
        A[1] = 0.5 * (a[2] * b[0] - a[0] * b[2]);
        A[2] = 0.5 * (a[0] * b[1] - a[1] * b[0]);
    } else {
        const double* x2 = q + 3 * (size_t)fv[2];
        const double* x3 = q + 3 * (size_t)fv[3];
        double a[3], b[3], c[3], d[3];
        for (int k = 0; k < 3; ++k) {
            a[k] = x1[k] - x0[k];
            b[k] = x3[k] - x0[k];
            c[k] = x3[k] - x2[k];
            d[k] = x1[k] - x2[k];
        }
        A[0] = 0.5 * (a[1] * b[2] - a[2] * b[1]);
        A[1] = 0.5 * (a[2] * b[0] - a[0] * b[2]);
        A[2] = 0.5 * (a[0] * b[1] - a[1] * b[0]);
        A[0] += 0.5 * (c[1] * d[2] - c[2] * d[1]);
        A[1] += 0.5 * (c[2] * d[0] - c[0] * d[2]);
        A[2] += 0.5 * (c[0] * d[1] - c[1] * d[0]);
    }
}

/* PressureForce::compute_force (pressureForce.cpp:14-55): -p A / nv on every
 * vertex of the face, faces in order */
void orc_pressure_force(const orc_sys* s, const double* q, double* f) {
    memset(f, 0, sizeof(double) * s->n);
    for (int fi = 0; fi < s->npf; ++fi) {
        double A[3];
        face_area(s, fi, q, A);
        const double p = s->pact[s->pgroup[fi]];
        double fP[3];
        for (int k = 0; k < 3; ++k) fP[k] = -p * A[k] / s->pnv;
        for (int j = 0; j < s->pnv; ++j)
            for (int k = 0; k < 3; ++k) f[3 * (size_t)s->pfaces[(size_t)s->pnv * fi + j] + k] += fP[k];
    }
}

/* the 3x3 matrix M(a, b) of the reference's hessX blocks: rows
 * (0, b2 - a2, a1 - b1), (a2 - b2, 0, b0 - a0), (b1 - a1, a0 - b0, 0) for
 * the block of vertex k with the other two triangle vertices a (next) and
 * b (previous) in the reference's order (pressureForce.cpp:66-80) */
static void cross_block(const double* xa, const double* xb, double* M) {
    M[0] = 0;
    M[1] = -xb[2] + xa[2];
    M[2] = xb[1] - xa[1];
    M[3] = xb[2] - xa[2];
    M[4] = 0;
    M[5] = -xb[0] + xa[0];
    M[6] = -xb[1] + xa[1];
    M[7] = xb[0] - xa[0];
    M[8] = 0;
}

/* PressureForce::compute_force_gradient (pressureForce.cpp:60-185): block
 * (j, k) = p / 6 * hessX[k], the same for every row vertex j.  Triangles:
 * hessX0 = M(x1, x2), hessX1 = M(x2, x0), hessX2 = M(x0, x1).  Quads: the
 * triangle (x0, x1, x3) gives hessX0, hessX1, hessX3 by the same rule and the
 * triangle (x2, x3, x1) adds hessX2 (set) and its x3 / x1 blocks (added). */
void orc_pressure_jacobian(const orc_sys* s, const double* q, double* J) {
    const int nv = s->pnv;
    for (int fi = 0; fi < s->npf; ++fi) {
        const int* fv = s->pfaces + (size_t)nv * fi;
        double H[4][9];
        if (nv == 3) {
            const double *x0 = q + 3 * (size_t)fv[0], *x1 = q + 3 * (size_t)fv[1], *x2 = q + 3 * (size_t)fv[2];
            cross_block(x1, x2, H[0]);
            cross_block(x2, x0, H[1]);
            cross_block(x0, x1, H[2]);
        } else {
            const double *q0 = q + 3 * (size_t)fv[0], *q1 = q + 3 * (size_t)fv[1], *q2 = q + 3 * (size_t)fv[2],
                         *q3 = q + 3 * (size_t)fv[3];
            double T[9];
            /* first triangle: x0 = q0, x1 = q1, x2 = q3 */
            cross_block(q1, q3, H[0]);
            cross_block(q3, q0, H[1]);
            cross_block(q0, q1, H[3]);
            /* second triangle: x0 = q2, x1 = q3, x2 = q1 */
            cross_block(q3, q1, H[2]);
            cross_block(q1, q2, T);     /* its x1 (= q3) block */
            for (int t = 0; t < 9; ++t) H[3][t] += T[t];
            cross_block(q2, q3, T);     /* its x2 (= q1) block */
            for (int t = 0; t < 9; ++t) H[1][t] += T[t];
        }
        const double p = s->pact[s->pgroup[fi]];
        for (int j = 0; j < nv; ++j)
            for (int k = 0; k < nv; ++k)
                for (int t = 0; t < 9; ++t) J[(((size_t)fi * nv + j) * nv + k) * 9 + t] = p / 6 * H[k][t];
    }
}

int orc_ndof(const orc_sys* s) { return s->n; }
void orc_get_mass(const orc_sys* s, double* m) { memcpy(m, s->mass, sizeof(double) * s->n); }
/* integrator.h:101-102 */
void orc_set_initial_velocity(orc_sys* s, const double* v) { memcpy(s->vprev, v, sizeof(double) * s->n); }
void orc_set_initial_deformation(orc_sys* s, const double* q) { memcpy(s->qprev, q, sizeof(double) * s->n); }
void orc_get_state(const orc_sys* s, double* qprev, double* vprev) {
    if (qprev) memcpy(qprev, s->qprev, sizeof(double) * s->n);
    if (vprev) memcpy(vprev, s->vprev, sizeof(double) * s->n);
}
void orc_set_threads(orc_sys* s, int nthreads) { s->nthreads = nthreads > 0 ? nthreads : 1; }

static void gather(const orc_sys* s, int e, const double* q, double* xe) {
    for (int a = 0; a < s->ed; ++a)
        for (int k = 0; k < 3; ++k) xe[3 * a + k] = q[3 * (size_t)s->conn[(size_t)e * s->ed + a] + k];
}

static int elem(const orc_sys* s, int e, const double* xe, int project, double* energy, double* grad,
                double* hess) {
    if (s->poly) {
        poly_eval(xe, energy, grad, hess);
        return ORC_OK;
    }
    int psz = (s->ed == 4) ? 9 : 192;
    double sa = 0.0;
    const double* f = NULL;
    if (s->agroup && s->agroup[e] >= 0) {
        sa = s->sigma * s->act[s->agroup[e]];
        f = s->afib + 3 * (size_t)e;
    }
    const double* md = NULL;
    double mk = 0.0, ma = 0.0;
    if (s->mgroup && s->mgroup[e] >= 0) {
        md = s->mdir + 3 * (size_t)e;
        mk = s->mk[e];
        ma = s->mact[s->mgroup[e]];
    }
    return elem_eval_pre(s->ed, s->pre + (size_t)e * psz, s->V0[e], xe, s->mu[e], s->lam[e], project, energy, grad,
                         hess, sa, f, s->material, md, mk, ma);
}

/* TimeIntegrator::integrate_energy (integrator.cpp:3-43) */
static double integrate_energy(const orc_sys* s, const double* q, double dt, double eint) {
    double grav = 0;
    for (int i = 0; i < s->n; ++i) grav -= s->mass[i] * q[i] * s->grav[i % 3];
    if (dt == 0) return grav + eint;
    double kin = 0, damp = 0;
    for (int i = 0; i < s->n; ++i) {
        double v, a;
        if (s->scheme == ORC_BE) {
            v = (q[i] - s->qprev[i]) / dt;
        } else {
            v = 2 * (q[i] - s->qprev[i]) / dt - s->vprev[i];
        }
        a = (v - s->vprev[i]) / dt;
        kin += s->mass[i] * a * a;
        damp += s->mass[i] * v * v;
    }
    if (s->scheme == ORC_BE) {
        kin = 0.5 * kin * dt * dt;
        damp = 0.5 * s->alpha * dt * damp;
    } else {
        kin = 0.5 * kin * dt * dt / 2;
        damp = 0.25 * s->alpha * dt * damp;
    }
    return kin + grav + damp + eint;
}


/* ---- contact -------------------------------------------------------------
 * Penalty, per candidate node x and primitive j (plane or disk):
 *   d = n.(x - p)  (plane point p / disk centre c), active when d < 0;
 *   E = 1/2 k w d^2,  w = 1 for a plane; for a disk w = s(rho) with
 *   t = (x - c) - d n, rho = |t|, s = 1 for rho <= r, 1 - 3z^2 + 2z^3 with
 *   z = (rho - r)/delta on the rim (delta = 1e-3 m), 0 beyond;
 *   gradient  k w d n + 1/2 k d^2 s' t/rho;
 *   Hessian   k w n n^T + k d s' (n th^T + th n^T)
 *             + 1/2 k d^2 (s'' th th^T + s' (I - n n^T - th th^T)/rho), th = t/rho.
 * Friction (lagged, incremental-potential form; the C2 smoothing f0 of Li et
 * al., "Incremental Potential Contact", SIGGRAPH 2020, Sec. 5): at the start
 * of a step the primitive with the largest penalty force k w |d| at qprev
 * gives the node its normal force N, normal n and the primitive's
 * displacement D over the step (disk: centre now minus centre at the end of
 * the last step).  With u = (x - qprev) - D, uT = (I - n n^T) u, y = |uT|,
 * eps = eps_v dt:
 *   F = mu N f0(y),  f0 = -y^3/(3 eps^2) + y^2/eps + eps/3 (y < eps), y (y >= eps);
 *   gradient mu N (f1(y)/y) uT, f1 = f0' = -y^2/eps^2 + 2y/eps (y < eps), 1;
 *   Hessian  mu N [(f1/y) P - uT uT^T/(y eps^2)]   (y < eps)
 *            mu N / y (P - uT uT^T/y^2)            (y >= eps),  P = I - n n^T.
 * The node's 3x3 contact block (all primitives + friction) is projected to
 * PSD (eigenvalues clamped at 0); it only differs from the exact block on a
 * disk rim.  Masked DOFs take no contact terms. */
static void normalize3(double* v) {
    double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (l > 0)
        for (int k = 0; k < 3; ++k) v[k] /= l;
}

int orc_set_contact(orc_sys* s, int nplanes, const double* pp, const double* pn, int ndisks, const double* dc,
                    const double* dn, const double* dr, int nsurf, const int* surf, double stiffness,
                    double friction, double eps_v) {
    if (nplanes < 0 || nplanes > ORC_MAX_PRIM || ndisks < 0 || ndisks > ORC_MAX_PRIM || stiffness < 0 ||
        friction < 0 || (friction > 0 && !(eps_v > 0)))
        return ORC_ERR_ARG;
    s->nplanes = nplanes;
    s->ndisks = ndisks;
    for (int j = 0; j < nplanes; ++j)
        for (int k = 0; k < 3; ++k) {
            s->plane_p[j][k] = pp[3 * j + k];
            s->plane_n[j][k] = pn[3 * j + k];
        }
    for (int j = 0; j < nplanes; ++j) normalize3(s->plane_n[j]); /* Plane3D(point, normal.normalized()) */
    for (int j = 0; j < ndisks; ++j) {
        for (int k = 0; k < 3; ++k) {
            s->disk_c[j][k] = s->disk_c_last[j][k] = dc[3 * j + k];
            s->disk_n[j][k] = dn[3 * j + k];
        }
        normalize3(s->disk_n[j]); /* Disk3D(center, normal.normalized(), radius) */
        s->disk_r[j] = dr[j];
    }
    free(s->surf);
    free(s->lagN);
    free(s->lagn);
    free(s->lagD);
    s->nsurf = surf ? nsurf : s->N;
    s->surf = (int*)malloc(sizeof(int) * (s->nsurf > 0 ? s->nsurf : 1));
    for (int i = 0; i < s->nsurf; ++i) s->surf[i] = surf ? surf[i] : i;
    s->lagN = (double*)calloc(s->nsurf > 0 ? s->nsurf : 1, sizeof(double));
    s->lagn = (double*)calloc(3 * (size_t)(s->nsurf > 0 ? s->nsurf : 1), sizeof(double));
    s->lagD = (double*)calloc(3 * (size_t)(s->nsurf > 0 ? s->nsurf : 1), sizeof(double));
    free(s->pmask);
    free(s->dmask);
    s->pmask = (int*)malloc(sizeof(int) * (s->nsurf > 0 ? s->nsurf : 1));
    s->dmask = (int*)malloc(sizeof(int) * (s->nsurf > 0 ? s->nsurf : 1));
    for (int i = 0; i < s->nsurf; ++i) s->pmask[i] = s->dmask[i] = -1;
    s->dfrozen = 0;
    s->kc = stiffness;
    s->muf = friction;
    s->eps_v = eps_v;
    s->lag_dt = 0;
    s->has_contact = (nplanes + ndisks) > 0;
    orc_contact_lag(s, 0.0);
    return ORC_OK;
}

int orc_set_disk(orc_sys* s, int idx, const double c[3], const double n[3], double r) {
    if (idx < 0 || idx >= s->ndisks) return ORC_ERR_ARG;
    for (int k = 0; k < 3; ++k) {
        s->disk_c[idx][k] = c[k];
        s->disk_n[idx][k] = n[k];
    }
    normalize3(s->disk_n[idx]);
    s->disk_r[idx] = r;
    return ORC_OK;
}

/* rim weight s(rho), ds/drho and d2s/drho2 */
static double rim_weight(double rho, double r, double* ds, double* d2s) {
    *ds = 0;
    *d2s = 0;
    if (rho <= r) return 1.0;
    double z = (rho - r) / ORC_DISK_RIM;
    if (z >= 1.0) return 0.0;
    *ds = (-6 * z + 6 * z * z) / ORC_DISK_RIM;
    *d2s = (-6 + 12 * z) / (ORC_DISK_RIM * ORC_DISK_RIM);
    return 1 - 3 * z * z + 2 * z * z * z;
}

/* penalty of one node against primitive j (j < nplanes: plane, else disk
 * j - nplanes) with disk centre cc: returns energy, adds gradient/Hessian */
static double penalty_term(const orc_sys* s, int j, const double* x, const double* cc, double* grad, double* H) {
    const double *p, *nn;
    if (j < s->nplanes) {
        p = s->plane_p[j];
        nn = s->plane_n[j];
    } else {
        p = cc;
        nn = s->disk_n[j - s->nplanes];
    }
    double v[3] = {x[0] - p[0], x[1] - p[1], x[2] - p[2]};
    double d = nn[0] * v[0] + nn[1] * v[1] + nn[2] * v[2];
    if (!(d < 0)) return 0.0;
    double w = 1.0, ds = 0.0, d2s = 0.0, th[3] = {0, 0, 0}, rho = 0;
    if (j >= s->nplanes) {
        double tv[3];
        for (int k = 0; k < 3; ++k) tv[k] = v[k] - d * nn[k];
        rho = sqrt(tv[0] * tv[0] + tv[1] * tv[1] + tv[2] * tv[2]);
        w = rim_weight(rho, s->disk_r[j - s->nplanes], &ds, &d2s);
        if (w == 0.0) return 0.0;
        if (rho > 0)
            for (int k = 0; k < 3; ++k) th[k] = tv[k] / rho;
        else
            ds = d2s = 0.0; /* on the axis: inside the radius */
    }
    double kc = s->kc;
    if (grad)
        for (int k = 0; k < 3; ++k) grad[k] += kc * w * d * nn[k] + 0.5 * kc * d * d * ds * th[k];
    if (H)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                double h = kc * w * nn[a] * nn[b];
                if (ds != 0.0 || d2s != 0.0) {
                    double P = (a == b) - nn[a] * nn[b] - th[a] * th[b];
                    h += kc * d * ds * (nn[a] * th[b] + th[a] * nn[b]) +
                         0.5 * kc * d * d * (d2s * th[a] * th[b] + ds * P / rho);
                }
                H[3 * a + b] += h;
            }
    return 0.5 * kc * w * d * d;
}

static double friction_term(const orc_sys* s, int i, const double* x, const double* xprev, double dt, double* grad,
                            double* H) {
    double N = s->lagN[i];
    if (!(N > 0) || !(s->muf > 0)) return 0.0;
    const double* nn = s->lagn + 3 * (size_t)i;
    const double* D = s->lagD + 3 * (size_t)i;
    double u[3], uT[3];
    for (int k = 0; k < 3; ++k) u[k] = x[k] - xprev[k] - D[k];
    double un = nn[0] * u[0] + nn[1] * u[1] + nn[2] * u[2];
    for (int k = 0; k < 3; ++k) uT[k] = u[k] - un * nn[k];
    double y = sqrt(uT[0] * uT[0] + uT[1] * uT[1] + uT[2] * uT[2]);
    double eps = s->eps_v * dt, muN = s->muf * N;
    double f1y, c2, f0; /* f1(y)/y and the coefficient of uT uT^T */
    if (y < eps) {
        f1y = -y / (eps * eps) + 2.0 / eps;
        c2 = y > 0 ? -1.0 / (y * eps * eps) : 0.0;
        f0 = -y * y * y / (3 * eps * eps) + y * y / eps + eps / 3;
    } else {
        f1y = 1.0 / y;
        c2 = -1.0 / (y * y * y);
        f0 = y;
    }
    if (grad)
        for (int k = 0; k < 3; ++k) grad[k] += muN * f1y * uT[k];
    if (H)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                H[3 * a + b] += muN * (f1y * ((a == b) - nn[a] * nn[b]) + c2 * uT[a] * uT[b]);
    return muN * f0;
}

/* symmetric 3x3: clamp negative eigenvalues to 0 (cyclic Jacobi) */
static void project_psd3(double* H) {
    orc_project_psd(3, H);
}

/* energy, and optionally gradient (added into g) / node blocks (into Hn,
 * [nsurf][9]) of all contact terms at q */
static double contact_eval(const orc_sys* s, const double* q, double dt, double* g, double* Hn) {
    /* under minimize_SQP the planes and disks are constraints, not energy */
    if (!s->has_contact || s->method == ORC_SQP) return 0.0;
    double e = 0;
    for (int i = 0; i < s->nsurf; ++i) {
        int node = s->surf[i];
        const double* x = q + 3 * (size_t)node;
        double gr[3] = {0, 0, 0}, H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int j = 0; j < s->nplanes + s->ndisks; ++j)
            e += penalty_term(s, j, x, j < s->nplanes ? NULL : s->disk_c[j - s->nplanes], gr, H);
        if (dt != 0) e += friction_term(s, i, x, s->qprev + 3 * (size_t)node, dt, gr, H);
        if (Hn) project_psd3(H);
        if (g)
            for (int k = 0; k < 3; ++k)
                if (!s->mask[3 * (size_t)node + k]) g[3 * (size_t)node + k] += gr[k];
        if (Hn) memcpy(Hn + 9 * (size_t)i, H, sizeof(H));
    }
    return e;
}

double orc_contact_energy(orc_sys* s, const double* q, double dt) { return contact_eval(s, q, dt, NULL, NULL); }

void orc_contact_lag(orc_sys* s, double dt) {
    s->lag_dt = dt;
    if (!s->has_contact) return;
    for (int i = 0; i < s->nsurf; ++i) {
        const double* x = s->qprev + 3 * (size_t)s->surf[i];
        double best = 0, bn[3] = {0, 0, 0}, bD[3] = {0, 0, 0};
        for (int j = 0; j < s->nplanes + s->ndisks; ++j) {
            double gr[3] = {0, 0, 0};
            const double* cc = j < s->nplanes ? NULL : s->disk_c_last[j - s->nplanes];
            penalty_term(s, j, x, cc, gr, NULL);
            const double* nn = j < s->nplanes ? s->plane_n[j] : s->disk_n[j - s->nplanes];
            double f = fabs(gr[0] * nn[0] + gr[1] * nn[1] + gr[2] * nn[2]); /* k w |d| */
            if (f > best) {
                best = f;
                for (int k = 0; k < 3; ++k) {
                    bn[k] = nn[k];
                    bD[k] = j < s->nplanes ? 0.0 : s->disk_c[j - s->nplanes][k] - s->disk_c_last[j - s->nplanes][k];
                }
            }
        }
        s->lagN[i] = best;
        for (int k = 0; k < 3; ++k) {
            s->lagn[3 * (size_t)i + k] = bn[k];
            s->lagD[3 * (size_t)i + k] = bD[k];
        }
    }
}

/* Energy::compute_energy (energy.cpp:208-248), external forces omitted */
double orc_energy(orc_sys* s, const double* q, double dt) {
    double energy = 0;
    int bad = 0;
#pragma omp parallel for reduction(+ : energy) reduction(| : bad) num_threads(s->nthreads) schedule(static)
    for (int e = 0; e < s->nel; ++e) {
        double xe[24], ee;
        gather(s, e, q, xe);
        elem(s, e, xe, 0, &ee, NULL, NULL);
        energy += ee;
    }
    (void)bad;
    if (s->fext)
        for (int i = 0; i < s->n; ++i) energy += s->fext[i] * (q[i] - s->qprev[i]);
    if (s->npf > 0) {
        double* fp = (double*)malloc(sizeof(double) * s->n);
        orc_pressure_force(s, q, fp);
        for (int i = 0; i < s->n; ++i) energy += fp[i] * (q[i] - s->qprev[i]);
        free(fp);
    }
    return integrate_energy(s, q, dt, energy) + contact_eval(s, q, dt, NULL, NULL);
}

/* Energy::compute_gradient (energy.cpp:252-308) + integrate_gradient
 * (integrator.cpp:45-84).  Masked element terms are dropped unless SQP
 * semantics (:275); the Dirichlet override g_i = q_i - qprev_i - dt*v_i is
 * applied for the Newton methods (:299-306). */
int orc_gradient(orc_sys* s, const double* q, double dt, double* g) {
    int n = s->n, nt = s->nthreads;
    double* acc = (double*)calloc((size_t)n * nt, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double* gl = acc + (size_t)tid * n;
#pragma omp for schedule(static)
        for (int e = 0; e < s->nel; ++e) {
            double xe[24], ge[24];
            gather(s, e, q, xe);
            elem(s, e, xe, 0, NULL, ge, NULL);
            for (int a = 0; a < s->ed; ++a)
                for (int k = 0; k < 3; ++k) {
                    int R = 3 * s->conn[(size_t)e * s->ed + a] + k;
                    if (s->method == ORC_SQP || s->mask[R] == 0) gl[R] += ge[3 * a + k];
                }
        }
    }
    for (int i = 0; i < n; ++i) {
        double v = 0;
        for (int t = 0; t < nt; ++t) v += acc[(size_t)t * n + i];
        g[i] = v;
    }
    free(acc);
    if (s->fext)
        for (int i = 0; i < n; ++i) g[i] -= s->fext[i];
    if (s->npf > 0) {
        double* fp = (double*)malloc(sizeof(double) * n);
        orc_pressure_force(s, q, fp);
        for (int i = 0; i < n; ++i) g[i] -= fp[i];
        free(fp);
    }
    /* integrate_gradient */
    if (dt != 0) {
        memcpy(s->gprevtmp, g, sizeof(double) * n);
        for (int i = 0; i < n; ++i) {
            double gg = -s->mass[i] * s->grav[i % 3];
            double v;
            if (s->scheme == ORC_BE) {
                v = (q[i] - s->qprev[i]) / dt;
                g[i] = s->mass[i] * ((v - s->vprev[i]) / dt) + gg + s->alpha * s->mass[i] * v + g[i];
            } else {
                v = 2 * (q[i] - s->qprev[i]) / dt - s->vprev[i];
                g[i] = s->mass[i] * ((v - s->vprev[i]) / dt) + gg + s->alpha * s->mass[i] * v +
                       0.5 * (g[i] + s->gprev[i]);
            }
        }
    } else {
        for (int i = 0; i < n; ++i) g[i] += -s->mass[i] * s->grav[i % 3];
    }
    contact_eval(s, q, dt, g, NULL);
    if (s->method != ORC_SQP)
        for (int i = 0; i < n; ++i)
            if (s->mask[i]) g[i] = -(s->qprev[i] + dt * s->bcval[i] - q[i]);
    return ORC_OK;
}

static int cmp_int(const void* a, const void* b) {
    int x = *(const int*)a, y = *(const int*)b;
    return (x > y) - (x < y);
}

/* CSR pattern of Energy::compute_hessian (energy.cpp:312-390) after
 * setFromTriplets (:377), integrate_hessian (integrator.cpp:87-124) and the
 * unit diagonal of masked DOFs (:381-388).  Rule:
 *   {(R,C): R,C DOFs of a common element, both unmasked (all under SQP)}
 *   u {(0,0) if any element pair was skipped: default Triplet() slots}
 *   u {(i,i) for all i if dt != 0} u {masked (i,i) for the Newton methods}.
 * Built per row from node adjacency; brute-force triplet enumeration checks
 * it in the tests. */
long long orc_pattern(orc_sys* s, double dt) {
    int dtnz = (dt != 0);
    if (s->pattern_dtnz == dtnz) return s->nnz;
    free(s->rowptr);
    free(s->cols);
    int N = s->N, n = s->n, ed = s->ed;
    /* node -> element incidence */
    int* cnt = (int*)calloc(N + 1, sizeof(int));
    for (size_t i = 0; i < (size_t)s->nel * ed; ++i) cnt[s->conn[i] + 1]++;
    for (int i = 0; i < N; ++i) cnt[i + 1] += cnt[i];
    int* inc = (int*)malloc(sizeof(int) * (size_t)s->nel * ed);
    int* fill = (int*)calloc(N, sizeof(int));
    for (int e = 0; e < s->nel; ++e)
        for (int a = 0; a < ed; ++a) {
            int v = s->conn[(size_t)e * ed + a];
            inc[cnt[v] + fill[v]++] = e;
        }
    int anyskip = 0;
    if (s->method != ORC_SQP)
        for (size_t i = 0; i < (size_t)s->nel * ed && !anyskip; ++i)
            for (int k = 0; k < 3; ++k)
                if (s->mask[3 * s->conn[i] + k]) anyskip = 1;
    /* per node sorted unique neighbours */
    int* nb_ptr = (int*)calloc(N + 1, sizeof(int));
    int cap = 0;
    for (int v = 0; v < N; ++v) cap = (cnt[v + 1] - cnt[v]) * ed > cap ? (cnt[v + 1] - cnt[v]) * ed : cap;
    int* tmp = (int*)malloc(sizeof(int) * (cap + 1));
    int* nbc = (int*)malloc(sizeof(int) * (size_t)(cnt[N]) * ed + 1);
    for (int v = 0; v < N; ++v) {
        int m = 0;
        for (int t = cnt[v]; t < cnt[v + 1]; ++t)
            for (int a = 0; a < ed; ++a) tmp[m++] = s->conn[(size_t)inc[t] * ed + a];
        qsort(tmp, m, sizeof(int), cmp_int);
        int u = 0;
        for (int i = 0; i < m; ++i)
            if (u == 0 || tmp[i] != tmp[u - 1]) tmp[u++] = tmp[i];
        nb_ptr[v + 1] = nb_ptr[v] + u;
        memcpy(nbc + nb_ptr[v], tmp, sizeof(int) * u);
    }
    /* rows: element-pair columns, then the forced diagonal inserted in order */
    s->rowptr = (int*)calloc(n + 1, sizeof(int));
    int sqp = (s->method == ORC_SQP);
    int* row = (int*)malloc(sizeof(int) * (3 * (size_t)cap + 2));
    for (int pass = 0; pass < 2; ++pass) {
        long long pos = 0;
        for (int R = 0; R < n; ++R) {
            int v = R / 3, m = 0;
            if (sqp || !s->mask[R])
                for (int t = nb_ptr[v]; t < nb_ptr[v + 1]; ++t)
                    for (int l = 0; l < 3; ++l) {
                        int C = 3 * nbc[t] + l;
                        if (sqp || !s->mask[C]) row[m++] = C;
                    }
            int diag_forced = dtnz || (!sqp && s->mask[R]) || (R == 0 && anyskip);
            if (diag_forced) {
                int at = 0;
                while (at < m && row[at] < R) ++at;
                if (at == m || row[at] != R) {
                    memmove(row + at + 1, row + at, sizeof(int) * (m - at));
                    row[at] = R;
                    ++m;
                }
            }
            if (pass)
                memcpy(s->cols + pos, row, sizeof(int) * m);
            else
                s->rowptr[R + 1] = m;
            pos += m;
        }
        if (!pass) {
            for (int R = 0; R < n; ++R) s->rowptr[R + 1] += s->rowptr[R];
            s->nnz = s->rowptr[n];
            s->cols = (int*)malloc(sizeof(int) * (size_t)(s->nnz > 0 ? s->nnz : 1));
        }
    }
    free(row);
    free(cnt);
    free(inc);
    free(fill);
    free(nb_ptr);
    free(tmp);
    free(nbc);
    s->pattern_dtnz = dtnz;
    return s->nnz;
}

void orc_get_pattern(const orc_sys* s, int* rowptr, int* cols) {
    memcpy(rowptr, s->rowptr, sizeof(int) * (s->n + 1));
    memcpy(cols, s->cols, sizeof(int) * (size_t)s->nnz);
}

static int csr_find(const orc_sys* s, int R, int C) {
    int lo = s->rowptr[R], hi = s->rowptr[R + 1] - 1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        if (s->cols[mid] == C) return mid;
        if (s->cols[mid] < C)
            lo = mid + 1;
        else
            hi = mid - 1;
    }
    return -1;
}

/* scatter map: for element e, local (r, c) -> CSR value index, -1 where the
 * reference skips the triplet (energy.cpp:337). */
void orc_get_scatter_map(const orc_sys* s, int* map) {
    int nd = s->nd, ed = s->ed;
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
    for (int e = 0; e < s->nel; ++e)
        for (int r = 0; r < nd; ++r)
            for (int c = 0; c < nd; ++c) {
                int R = 3 * s->conn[(size_t)e * ed + r / 3] + r % 3;
                int C = 3 * s->conn[(size_t)e * ed + c / 3] + c % 3;
                int keep = (s->method == ORC_SQP) || (!s->mask[R] && !s->mask[C]);
                map[((size_t)e * nd + r) * nd + c] = keep ? csr_find(s, R, C) : -1;
            }
}

/* Hessian values on the pattern (energy.cpp:312-390, integrator.cpp:87-124). */
int orc_hessian(orc_sys* s, const double* q, double dt, int project, double* vals) {
    orc_pattern(s, dt);
    int nd = s->nd, ed = s->ed;
    memset(vals, 0, sizeof(double) * (size_t)s->nnz);
    int status = ORC_OK;
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
    for (int e = 0; e < s->nel; ++e) {
        double xe[24], he[576];
        gather(s, e, q, xe);
        int st = elem(s, e, xe, project, NULL, NULL, he);
        if (st != ORC_OK) {
#pragma omp atomic write
            status = st;
        }
        for (int r = 0; r < nd; ++r) {
            int R = 3 * s->conn[(size_t)e * ed + r / 3] + r % 3;
            for (int c = 0; c < nd; ++c) {
                int C = 3 * s->conn[(size_t)e * ed + c / 3] + c % 3;
                if (s->method != ORC_SQP && (s->mask[R] || s->mask[C])) continue;
                int idx = csr_find(s, R, C);
                if (s->nthreads > 1) {
#pragma omp atomic
                    vals[idx] += he[r * nd + c];
                } else {
                    vals[idx] += he[r * nd + c];
                }
            }
        }
    }
    if (s->npf > 0 && s->method == ORC_NEWTON) {
        /* force Jacobian triplets, unmasked rows and columns (energy.cpp:354-372) */
        const int nv = s->pnv;
        double* J = (double*)malloc(sizeof(double) * 9 * (size_t)s->npf * nv * nv);
        orc_pressure_jacobian(s, q, J);
        for (int fi = 0; fi < s->npf; ++fi)
            for (int j = 0; j < nv; ++j)
                for (int k = 0; k < nv; ++k)
                    for (int l = 0; l < 3; ++l)
                        for (int m = 0; m < 3; ++m) {
                            const int R = 3 * s->pfaces[(size_t)nv * fi + j] + l;
                            const int C = 3 * s->pfaces[(size_t)nv * fi + k] + m;
                            if (s->mask[R] || s->mask[C]) continue;
                            const int idx = csr_find(s, R, C);
                            if (idx < 0) {
                                status = ORC_ERR_ARG;  /* a face pair outside the element pattern */
                                continue;
                            }
                            vals[idx] += J[(((size_t)fi * nv + j) * nv + k) * 9 + 3 * l + m];
                        }
        free(J);
    }
    if (dt != 0) {
        double fac = (s->scheme == ORC_BE) ? 1.0 : 0.5;
        if (fac != 1.0)
            for (long long i = 0; i < s->nnz; ++i) vals[i] *= fac;
        for (int i = 0; i < s->n; ++i) {
            int idx = csr_find(s, i, i);
            double m = s->mass[i];
            vals[idx] += (s->scheme == ORC_BE) ? (m / (dt * dt) + s->alpha * m / dt)
                                               : (2 * m / (dt * dt) + 2 * s->alpha * m / dt);
        }
    }
    if (s->has_contact && s->method != ORC_SQP) {
        double* Hn = (double*)malloc(sizeof(double) * 9 * (size_t)(s->nsurf > 0 ? s->nsurf : 1));
        contact_eval(s, q, dt, NULL, Hn);
        for (int i = 0; i < s->nsurf; ++i)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    int R = 3 * s->surf[i] + a, C = 3 * s->surf[i] + b;
                    if (s->mask[R] || s->mask[C]) continue;
                    double h = Hn[9 * (size_t)i + 3 * a + b];
                    int idx = csr_find(s, R, C);
                    if (h != 0.0 && idx >= 0) vals[idx] += h;
                }
        free(Hn);
    }
    if (s->method != ORC_SQP)
        for (int i = 0; i < s->n; ++i)
            if (s->mask[i]) vals[csr_find(s, i, i)] = 1.0;
    return status;
}

/* ------------------------------------------------------------------------ */
/* Linear solvers                                                            */
/* ------------------------------------------------------------------------ */

/* Envelope (skyline) LDL^T without pivoting -- stands in for Eigen's
 * SimplicialLDLT (solver.cpp:65,82-84).  The factor is kept for repeated
 * solves (the QP solver below).  ldl_factor returns ORC_ERR_SOLVE on a zero
 * pivot. */
typedef struct {
    int n;
    int* first;      /* first column of row i's envelope */
    long long* off;  /* row i's envelope in L */
    double* L;
    double* D;
} ldl_fac;

static void ldl_free(ldl_fac* f) {
    free(f->first);
    free(f->off);
    free(f->L);
    free(f->D);
    memset(f, 0, sizeof(*f));
}

static int ldl_factor(int n, const int* rowptr, const int* cols, const double* vals, ldl_fac* f) {
    f->n = n;
    f->first = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
    f->off = (long long*)malloc(sizeof(long long) * (n + 1));
    for (int i = 0; i < n; ++i) {
        int fi = i;
        for (int p = rowptr[i]; p < rowptr[i + 1]; ++p)
            if (cols[p] < fi) fi = cols[p];
        f->first[i] = fi;
    }
    f->off[0] = 0;
    for (int i = 0; i < n; ++i) f->off[i + 1] = f->off[i] + (i - f->first[i]);
    f->L = (double*)calloc((size_t)f->off[n] + 1, sizeof(double));
    f->D = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
    double* u = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
    int status = ORC_OK;
    for (int i = 0; i < n && status == ORC_OK; ++i) {
        int fi = f->first[i];
        double aii = 0;
        for (int j = fi; j < i; ++j) u[j] = 0;
        for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            int c = cols[p];
            if (c < i) u[c] = vals[p];
            if (c == i) aii = vals[p];
        }
        double* Li = f->L + f->off[i] - fi; /* Li[j] for j in [fi, i) */
        for (int j = fi; j < i; ++j) {
            int fj = f->first[j];
            int k0 = fi > fj ? fi : fj;
            const double* Lj = f->L + f->off[j] - fj;
            double sum = u[j];
            for (int k = k0; k < j; ++k) sum -= u[k] * Lj[k];
            u[j] = sum;
        }
        double di = aii;
        for (int j = fi; j < i; ++j) {
            Li[j] = u[j] / f->D[j];
            di -= u[j] * Li[j];
        }
        if (di == 0 || di != di) status = ORC_ERR_SOLVE;
        f->D[i] = di;
    }
    free(u);
    if (status != ORC_OK) ldl_free(f);
    return status;
}

static void ldl_apply(const ldl_fac* f, const double* b, double* x) {
    const int n = f->n;
    for (int i = 0; i < n; ++i) x[i] = b[i];
    for (int i = 0; i < n; ++i) {
        const double* Li = f->L + f->off[i] - f->first[i];
        double sum = x[i];
        for (int j = f->first[i]; j < i; ++j) sum -= Li[j] * x[j];
        x[i] = sum;
    }
    for (int i = 0; i < n; ++i) x[i] /= f->D[i];
    for (int i = n - 1; i >= 0; --i) {
        const double* Li = f->L + f->off[i] - f->first[i];
        double xi = x[i];
        for (int j = f->first[i]; j < i; ++j) x[j] -= Li[j] * xi;
    }
}

static int ldlt_solve(int n, const int* rowptr, const int* cols, const double* vals, const double* b, double* x) {
    ldl_fac f;
    int status = ldl_factor(n, rowptr, cols, vals, &f);
    if (status != ORC_OK) return status;
    ldl_apply(&f, b, x);
    ldl_free(&f);
    return ORC_OK;
}

/* FP64 PCG with a 3x3 block-Jacobi preconditioner (the CPU form of the
 * device's linear solver).  Returns the iteration count or a negative code. */
static long long pcg_solve(int n, const int* rowptr, const int* cols, const double* vals, const double* b, double* x,
                           double rtol, int maxit, int nt) {
    int nb = n / 3;
    double* Minv = (double*)calloc((size_t)nb * 9, sizeof(double));
    for (int v = 0; v < nb; ++v) {
        double Bk[9] = {0};
        for (int k = 0; k < 3; ++k) {
            int R = 3 * v + k;
            for (int p = rowptr[R]; p < rowptr[R + 1]; ++p) {
                int C = cols[p];
                if (C >= 3 * v && C < 3 * v + 3) Bk[3 * k + (C - 3 * v)] = vals[p];
            }
        }
        double d = det3(Bk);
        if (d != 0 && d == d) {
            inv3(Bk, Minv + 9 * (size_t)v);
        } else {
            for (int k = 0; k < 3; ++k) Minv[9 * (size_t)v + 4 * k] = Bk[4 * k] != 0 ? 1.0 / Bk[4 * k] : 1.0;
        }
    }
    double *r = (double*)malloc(sizeof(double) * n), *z = (double*)malloc(sizeof(double) * n),
           *p = (double*)malloc(sizeof(double) * n), *Ap = (double*)malloc(sizeof(double) * n);
    for (int i = 0; i < n; ++i) {
        x[i] = 0;
        r[i] = b[i];
    }
    double bb = 0;
#pragma omp parallel for reduction(+ : bb) num_threads(nt) schedule(static)
    for (int i = 0; i < n; ++i) bb += b[i] * b[i];
    long long it = 0;
    double rz = 0;
    if (bb == 0) goto done;
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int v = 0; v < nb; ++v)
        for (int k = 0; k < 3; ++k) {
            const double* M = Minv + 9 * (size_t)v + 3 * k;
            z[3 * v + k] = M[0] * r[3 * v] + M[1] * r[3 * v + 1] + M[2] * r[3 * v + 2];
        }
#pragma omp parallel for reduction(+ : rz) num_threads(nt) schedule(static)
    for (int i = 0; i < n; ++i) rz += r[i] * z[i];
    memcpy(p, z, sizeof(double) * n);
    for (it = 0; it < maxit;) {
        double pAp = 0;
#pragma omp parallel for reduction(+ : pAp) num_threads(nt) schedule(static)
        for (int i = 0; i < n; ++i) {
            double acc = 0;
            for (int q = rowptr[i]; q < rowptr[i + 1]; ++q) acc += vals[q] * p[cols[q]];
            Ap[i] = acc;
            pAp += acc * p[i];
        }
        if (!(pAp > 0)) {
            it = -1;
            goto done;
        }
        double alpha = rz / pAp;
        double rr = 0;
#pragma omp parallel for reduction(+ : rr) num_threads(nt) schedule(static)
        for (int i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * Ap[i];
            rr += r[i] * r[i];
        }
        ++it;
        if (rr <= rtol * rtol * bb) break;
        double rzn = 0;
#pragma omp parallel for reduction(+ : rzn) num_threads(nt) schedule(static)
        for (int v = 0; v < nb; ++v)
            for (int k = 0; k < 3; ++k) {
                const double* M = Minv + 9 * (size_t)v + 3 * k;
                double zz = M[0] * r[3 * v] + M[1] * r[3 * v + 1] + M[2] * r[3 * v + 2];
                z[3 * v + k] = zz;
                rzn += r[3 * v + k] * zz;
            }
        double beta = rzn / rz;
        rz = rzn;
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
done:
    free(Minv);
    free(r);
    free(z);
    free(p);
    free(Ap);
    return it;
}

static double norm2(int n, const double* v) {
    double s = 0;
    for (int i = 0; i < n; ++i) s += v[i] * v[i];
    return sqrt(s);
}

/* TimeIntegrator::update_state (integrator.h:79-97) via Energy::update_state
 * (energy.cpp:539-557) */
void orc_update_state(orc_sys* s, const double* q, double dt) {
    for (int i = 0; i < s->n; ++i) {
        double v = (s->scheme == ORC_BE) ? (q[i] - s->qprev[i]) / dt : 2 * (q[i] - s->qprev[i]) / dt - s->vprev[i];
        s->vprev[i] = v;
        s->qprev[i] = q[i];
        s->gprev[i] = s->gprevtmp[i];
    }
}

/* minimize_frozenhessian (solver.cpp:268-294): the Hessian of the initial
 * guess serves every iteration; x += H0^-1 (-g(x)) until ||dx|| < dx_tol (or
 * the optional residual criterion); no line search; running out of
 * iterations returns the iterate without an error, as the reference does. */
static int orc_minimize_frozen(orc_sys* s, double* x, double dt, const orc_opts* o, orc_stats* st) {
    int n = s->n;
    orc_pattern(s, dt);
    double* g = (double*)malloc(sizeof(double) * n);
    double* dx = (double*)malloc(sizeof(double) * n);
    double* rhs = (double*)malloc(sizeof(double) * n);
    double* vals = (double*)malloc(sizeof(double) * (size_t)(s->nnz > 0 ? s->nnz : 1));
    int status = orc_hessian(s, x, dt, o->project, vals);
    for (int it = 0; status == ORC_OK && it < o->max_iter; ++it) {
        orc_gradient(s, x, dt, g);
        st->gnorm = norm2(n, g);
        if (o->g_tol > 0 && st->gnorm < o->g_tol) break;
        for (int i = 0; i < n; ++i) rhs[i] = -g[i];
        if (o->linear_solver == ORC_LIN_LDLT) {
            if (ldlt_solve(n, s->rowptr, s->cols, vals, rhs, dx) != ORC_OK) {
                for (int i = 0; i < n; ++i) dx[i] = NAN;
            }
        } else {
            long long k = pcg_solve(n, s->rowptr, s->cols, vals, rhs, dx, o->pcg_rtol, o->pcg_max_iter, s->nthreads);
            if (k < 0) {
                status = ORC_ERR_SOLVE;
                break;
            }
            st->pcg_iters += k;
        }
        st->newton_iters++;
        double dxn = norm2(n, dx);
        if (dxn != dxn) {
            status = ORC_ERR_NAN;
            break;
        }
        for (int i = 0; i < n; ++i) x[i] += dx[i];
        st->dxnorm = dxn;
        if (dxn < o->dx_tol) break;
    }
    free(g);
    free(dx);
    free(rhs);
    free(vals);
    return status;
}

/* Newton minimisation: minimize_newton (solver.cpp:56-105) and
 * minimize_newton_linesearch (:201-264, Armijo alpha0 = 1, beta = 0.5,
 * c = 0.1) with the minimize_newton Dirichlet handling. */
static int orc_minimize_sqp(orc_sys* s, double* x, double dt, const orc_opts* o, orc_stats* st);

int orc_minimize(orc_sys* s, double* x, double dt, const orc_opts* o, orc_stats* st) {
    if (s->method == ORC_FROZEN) return orc_minimize_frozen(s, x, dt, o, st);
    if (s->method == ORC_SQP) return orc_minimize_sqp(s, x, dt, o, st);
    int n = s->n;
    orc_pattern(s, dt);
    double* g = (double*)malloc(sizeof(double) * n);
    double* dx = (double*)malloc(sizeof(double) * n);
    double* rhs = (double*)malloc(sizeof(double) * n);
    double* xn = (double*)malloc(sizeof(double) * n);
    double* vals = (double*)malloc(sizeof(double) * (size_t)(s->nnz > 0 ? s->nnz : 1));
    int status = ORC_ERR_NOCONV;
    int converged = 0;
    for (int it = 0; it < o->max_iter; ++it) {
        double t0 = now_s();
        orc_gradient(s, x, dt, g);
        double t1 = now_s();
        st->t_grad += t1 - t0;
        st->gnorm = norm2(n, g);
        if (o->g_tol > 0 && st->gnorm < o->g_tol) {
            converged = 1;
            break;
        }
        int hst = orc_hessian(s, x, dt, o->project, vals);
        double t2 = now_s();
        st->t_hess += t2 - t1;
        if (hst != ORC_OK) {
            status = hst;
            break;
        }
        for (int i = 0; i < n; ++i) rhs[i] = -g[i];
        if (o->linear_solver == ORC_LIN_LDLT) {
            if (ldlt_solve(n, s->rowptr, s->cols, vals, rhs, dx) != ORC_OK) {
                for (int i = 0; i < n; ++i) dx[i] = NAN;
            }
        } else {
            long long k = pcg_solve(n, s->rowptr, s->cols, vals, rhs, dx, o->pcg_rtol, o->pcg_max_iter, s->nthreads);
            if (k < 0) {
                status = ORC_ERR_SOLVE;
                break;
            }
            st->pcg_iters += k;
        }
        st->t_solve += now_s() - t2;
        st->newton_iters++;
        double dxn = norm2(n, dx);
        if (dxn != dxn) {
            status = ORC_ERR_NAN;
            break;
        }
        if (s->method != ORC_NEWTON_LS || dxn < o->dx_tol) {
            for (int i = 0; i < n; ++i) x[i] += dx[i];
            st->dxnorm = dxn;
            if (dxn < o->dx_tol) {
                converged = 1;
                break;
            }
            continue;
        }
        double te = now_s();
        double E0 = orc_energy(s, x, dt);
        double gdx = 0;
        for (int i = 0; i < n; ++i) gdx += g[i] * dx[i];
        double alpha = 1.0;
        int accepted = 0;
        for (int bt = 0; bt <= o->max_backtrack; ++bt) {
            for (int i = 0; i < n; ++i) xn[i] = x[i] + alpha * dx[i];
            double En = orc_energy(s, xn, dt);
            if (En <= E0 + 0.1 * alpha * gdx) {
                accepted = 1;
                break;
            }
            alpha *= 0.5;
            st->backtracks++;
        }
        st->t_energy += now_s() - te;
        if (!accepted) {
            if (alpha * 2 * dxn < o->dx_tol) {
                accepted = 1;
            } else {
                status = ORC_ERR_LS;
                break;
            }
        }
        double stepn = 0;
        for (int i = 0; i < n; ++i) {
            double d = xn[i] - x[i];
            stepn += d * d;
            x[i] = xn[i];
        }
        st->dxnorm = sqrt(stepn);
        if (st->dxnorm < o->dx_tol) {
            converged = 1;
            break;
        }
    }
    if (converged) status = ORC_OK;
    free(g);
    free(dx);
    free(rhs);
    free(xn);
    free(vals);
    return status;
}

/* ===========================================================================
 * minimize_SQP (solver.cpp:109-197) and its quadratic programs.
 *
 * Each SQP iteration of the reference hands the QP
 *     min 1/2 dx^T H dx + g^T dx   s.t.  l <= A dx <= u
 * to OSQP (default settings with scaling = 0 and max_iter = MAX_ITER_QP =
 * 250, solver.cpp:117-122; common.h:23).  OSQP is a dependency absent from
 * the reference tree (the empty submodule external/osqp, v1.0 API per the
 * OSQPSolver / OSQPCscMatrix_new calls of solver.cpp:147-163).  Its published
 * algorithm is restated here: the ADMM iteration of Stellato, Banjac,
 * Goulart, Bemporad and Boyd, "OSQP: an operator splitting solver for
 * quadratic programs", Math. Prog. Comp. 12 (2020), with its defaults
 * sigma = 1e-6, alpha = 1.6, rho = 0.1 (rho x 1e3 on equality rows), the
 * adaptive rho rule, and its solution polishing.  It is run until the
 * polished solution passes the KKT checks rather than stopped at eps = 1e-3
 * or 250 iterations: a strictly convex QP has one solution, which OSQP
 * approximates and at which the SQP iteration stops.
 *
 * Every constraint row of the reference acts on one vertex (NeumannBC: a
 * unit row per masked DOF; PlaneContact / DiskContact: the normal on a
 * surface vertex), so a row is (vertex, 3 coefficients, l, u).
 * ===========================================================================
 */
#define ORC_PLANE_THRESHOLD 0.01        /* PLANE_CONSTRAINT_THRESHOLD, common.h:29 */
#define ORC_DISK_PLANE_THRESHOLD 0.01   /* DISK_PLANAR_CONSTRAINT_THRESHOLD, common.h:33 */
#define ORC_DISK_RADIAL_THRESHOLD 0.001 /* DISK_RADIAL_CONSTRAINT_THRESHOLD, common.h:35 */

static double vmax_abs(int n, const double* v) {
    double m = 0;
    for (int i = 0; i < n; ++i) {
        double a = fabs(v[i]);
        if (a > m) m = a;
    }
    return m;
}

static int csr_at(const int* rowptr, const int* cols, int R, int C) {
    for (int p = rowptr[R]; p < rowptr[R + 1]; ++p)
        if (cols[p] == C) return p;
    return -1;
}

static void csr_mv(int n, const int* rowptr, const int* cols, const double* vals, const double* x, double* out) {
    for (int i = 0; i < n; ++i) {
        double acc = 0;
        for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) acc += vals[p] * x[cols[p]];
        out[i] = acc;
    }
}

/* out = A x */
static void rows_mv(int m, const int* vtx, const double* a, const double* x, double* out) {
    for (int i = 0; i < m; ++i) {
        const double* xv = x + 3 * (size_t)vtx[i];
        out[i] = a[3 * i] * xv[0] + a[3 * i + 1] * xv[1] + a[3 * i + 2] * xv[2];
    }
}

/* out += A^T y */
static void rows_mtv(int m, const int* vtx, const double* a, const double* y, double* out) {
    for (int i = 0; i < m; ++i)
        for (int k = 0; k < 3; ++k) out[3 * (size_t)vtx[i] + k] += a[3 * i + k] * y[i];
}

/* P + sigma I + A^T diag(rho) A on P's pattern (a row's a a^T lands in its
 * vertex's diagonal 3x3 block), factorised */
static int admm_factor(int n, const int* rowptr, const int* cols, const double* P, double sigma, int m,
                       const int* vtx, const double* a, const double* rho, double* work, ldl_fac* f) {
    memcpy(work, P, sizeof(double) * (size_t)rowptr[n]);
    for (int i = 0; i < n; ++i) {
        int d = csr_at(rowptr, cols, i, i);
        if (d < 0) return ORC_ERR_ARG;
        work[d] += sigma;
    }
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) {
                double v = rho[i] * a[3 * i + j] * a[3 * i + k];
                if (v == 0.0) continue;
                int p = csr_at(rowptr, cols, 3 * vtx[i] + j, 3 * vtx[i] + k);
                if (p < 0) return ORC_ERR_ARG;
                work[p] += v;
            }
    return ldl_factor(n, rowptr, cols, work, f);
}

/* OSQP's solution polishing: the QP restricted to the rows guessed active
 * from the ADMM iterate (lower: z_i - l_i < -y_i, upper: u_i - z_i < y_i;
 * equality rows always) is solved as the quasi-definite system
 *   [P + d I, A_a^T; A_a, -d I] [x; y_a] = [-q; b_a]    (d = 1e-6)
 * with iterative refinement against the same system without d.  Accepted
 * (returns 1) when A x lies within [l, u] and each active row's multiplier
 * has its bound's sign (y <= 0 on a lower bound, >= 0 on an upper one). */
static int qp_polish(int n, const int* rowptr, const int* cols, const double* P, const double* q, int m,
                     const int* vtx, const double* a, const double* l, const double* u, const double* z,
                     const double* yk, double* xo, double* yo) {
    const double delta = 1e-6;
    int* side = (int*)calloc(m > 0 ? m : 1, sizeof(int)); /* -1 lower, +1 upper, 2 equality, 0 inactive */
    int* aidx = (int*)malloc(sizeof(int) * (m > 0 ? m : 1));
    int ma = 0;
    for (int i = 0; i < m; ++i) {
        if (l[i] == u[i])
            side[i] = 2;
        else if (z[i] - l[i] < -yk[i])
            side[i] = -1;
        else if (u[i] - z[i] < yk[i])
            side[i] = 1;
        if (side[i]) aidx[ma++] = i;
    }
    const int N2 = n + ma;
    /* active rows per vertex, for the coupling columns of rows R < n */
    const int nv = n / 3;
    int* vcnt = (int*)calloc(nv + 1, sizeof(int));
    for (int j = 0; j < ma; ++j) vcnt[vtx[aidx[j]] + 1]++;
    for (int v = 0; v < nv; ++v) vcnt[v + 1] += vcnt[v];
    int* vrows = (int*)malloc(sizeof(int) * (ma > 0 ? ma : 1));
    int* vfill = (int*)malloc(sizeof(int) * (nv > 0 ? nv : 1));
    for (int v = 0; v < nv; ++v) vfill[v] = vcnt[v];
    for (int j = 0; j < ma; ++j) vrows[vfill[vtx[aidx[j]]]++] = j;
    long long nnz = rowptr[n] + 2LL * 3 * ma + ma;
    int* kp = (int*)malloc(sizeof(int) * (N2 + 1));
    int* kc = (int*)malloc(sizeof(int) * nnz);
    double* kv = (double*)malloc(sizeof(double) * nnz);   /* regularised */
    double* kh = (double*)malloc(sizeof(double) * nnz);   /* exact */
    long long t = 0;
    for (int R = 0; R < n; ++R) {
        kp[R] = (int)t;
        for (int p = rowptr[R]; p < rowptr[R + 1]; ++p) {
            kc[t] = cols[p];
            kh[t] = P[p];
            kv[t] = P[p] + (cols[p] == R ? delta : 0.0);
            ++t;
        }
        int v = R / 3, k = R % 3;
        for (int w = vcnt[v]; w < vcnt[v + 1]; ++w) {
            int j = vrows[w];
            kc[t] = n + j;
            kh[t] = kv[t] = a[3 * aidx[j] + k];
            ++t;
        }
    }
    for (int j = 0; j < ma; ++j) {
        int i = aidx[j];
        kp[n + j] = (int)t;
        for (int k = 0; k < 3; ++k) {
            kc[t] = 3 * vtx[i] + k;
            kh[t] = kv[t] = a[3 * i + k];
            ++t;
        }
        kc[t] = n + j;
        kh[t] = 0.0;
        kv[t] = -delta;
        ++t;
    }
    kp[N2] = (int)t;
    double* rhs = (double*)malloc(sizeof(double) * N2);
    double* sol = (double*)calloc(N2, sizeof(double));
    double* res = (double*)malloc(sizeof(double) * N2);
    double* cor = (double*)malloc(sizeof(double) * N2);
    for (int i = 0; i < n; ++i) rhs[i] = -q[i];
    for (int j = 0; j < ma; ++j) {
        int i = aidx[j];
        rhs[n + j] = side[i] == 1 ? u[i] : l[i];
    }
    ldl_fac f;
    int ok = ldl_factor(N2, kp, kc, kv, &f) == ORC_OK;
    if (ok) {
        ldl_apply(&f, rhs, sol);
        for (int it = 0; it < 25; ++it) {
            csr_mv(N2, kp, kc, kh, sol, res);
            for (int i = 0; i < N2; ++i) res[i] = rhs[i] - res[i];
            if (vmax_abs(N2, res) <= 1e-15 * (1.0 + vmax_abs(N2, rhs))) break;
            ldl_apply(&f, res, cor);
            for (int i = 0; i < N2; ++i) sol[i] += cor[i];
        }
        ldl_free(&f);
    }
    if (ok) {
        double* Ax = (double*)malloc(sizeof(double) * (m > 0 ? m : 1));
        double* Px = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
        rows_mv(m, vtx, a, sol, Ax);
        csr_mv(n, rowptr, cols, P, sol, Px);
        double bmax = 0;
        for (int i = 0; i < m; ++i) {
            if (fabs(Ax[i]) > bmax) bmax = fabs(Ax[i]);
            if (isfinite(l[i]) && fabs(l[i]) > bmax) bmax = fabs(l[i]);
        }
        const double tolp = 1e-13 + 1e-10 * bmax;
        const double told = 1e-13 + 1e-9 * (vmax_abs(n, q) > vmax_abs(n, Px) ? vmax_abs(n, q) : vmax_abs(n, Px));
        for (int i = 0; i < m && ok; ++i)
            if (Ax[i] < l[i] - tolp || Ax[i] > u[i] + tolp) ok = 0;
        for (int j = 0; j < ma && ok; ++j) {
            int i = aidx[j];
            double yj = sol[n + j];
            if ((side[i] == -1 && yj > told) || (side[i] == 1 && yj < -told)) ok = 0;
        }
        if (ok) {
            memcpy(xo, sol, sizeof(double) * n);
            for (int i = 0; i < m; ++i) yo[i] = 0.0;
            for (int j = 0; j < ma; ++j) yo[aidx[j]] = sol[n + j];
        }
        free(Ax);
        free(Px);
    }
    free(side);
    free(aidx);
    free(vcnt);
    free(vrows);
    free(vfill);
    free(kp);
    free(kc);
    free(kv);
    free(kh);
    free(rhs);
    free(sol);
    free(res);
    free(cor);
    return ok;
}

int orc_qp_solve(int n, const int* rowptr, const int* cols, const double* P, const double* q, int m, const int* vtx,
                 const double* a, const double* l, const double* u, double* x, double* y, int* iters) {
    const double sigma = 1e-6, alpha = 1.6;
    double rho = 0.1;
    int status = ORC_ERR_NOCONV;
    if (iters) *iters = 0;
    for (int i = 0; i < m; ++i)
        if (vtx[i] < 0 || 3 * vtx[i] + 2 >= n || !(l[i] <= u[i])) return ORC_ERR_ARG;
    double* work = (double*)malloc(sizeof(double) * (size_t)(rowptr[n] > 0 ? rowptr[n] : 1));
    if (m == 0) {
        /* unconstrained: P x = -q */
        double* b = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
        for (int i = 0; i < n; ++i) b[i] = -q[i];
        status = ldlt_solve(n, rowptr, cols, P, b, x);
        free(b);
        free(work);
        return status;
    }
    double* rv = (double*)malloc(sizeof(double) * m);     /* rho per row */
    double* z = (double*)calloc(m, sizeof(double));
    double* zp = (double*)malloc(sizeof(double) * m);
    double* zt = (double*)malloc(sizeof(double) * m);
    double* Ax = (double*)malloc(sizeof(double) * m);
    double* xt = (double*)malloc(sizeof(double) * n);
    double* rhs = (double*)malloc(sizeof(double) * n);
    double* Px = (double*)malloc(sizeof(double) * n);
    double* Aty = (double*)malloc(sizeof(double) * n);
    double* xs = (double*)malloc(sizeof(double) * n);
    double* ys = (double*)malloc(sizeof(double) * m);
    for (int i = 0; i < n; ++i) x[i] = 0.0;
    for (int i = 0; i < m; ++i) y[i] = 0.0;
    for (int i = 0; i < m; ++i) rv[i] = (l[i] == u[i]) ? 1e3 * rho : rho;
    ldl_fac f;
    int fst = admm_factor(n, rowptr, cols, P, sigma, m, vtx, a, rv, work, &f);
    if (fst != ORC_OK) status = fst;
    double eps = 1e-6;
    const int max_total = 400000;
    for (int k = 0; fst == ORC_OK && k < max_total; ++k) {
        /* (P + sigma I + A^T R A) xt = sigma x - q + A^T (R z - y) */
        for (int i = 0; i < n; ++i) rhs[i] = sigma * x[i] - q[i];
        for (int i = 0; i < m; ++i) zp[i] = rv[i] * z[i] - y[i];
        rows_mtv(m, vtx, a, zp, rhs);
        ldl_apply(&f, rhs, xt);
        rows_mv(m, vtx, a, xt, zt);
        for (int i = 0; i < n; ++i) x[i] = alpha * xt[i] + (1 - alpha) * x[i];
        for (int i = 0; i < m; ++i) {
            zp[i] = z[i];
            double w = alpha * zt[i] + (1 - alpha) * zp[i];
            double zn = w + y[i] / rv[i];
            z[i] = zn < l[i] ? l[i] : (zn > u[i] ? u[i] : zn);
            y[i] += rv[i] * (w - z[i]);
        }
        if (iters) *iters = k + 1;
        if ((k + 1) % 25) continue;
        /* residuals (termination relative to the scales sp, sd at eps) */
        rows_mv(m, vtx, a, x, Ax);
        csr_mv(n, rowptr, cols, P, x, Px);
        for (int i = 0; i < n; ++i) Aty[i] = 0.0;
        rows_mtv(m, vtx, a, y, Aty);
        double rp = 0, rd = 0;
        for (int i = 0; i < m; ++i) rp = fmax(rp, fabs(Ax[i] - z[i]));
        for (int i = 0; i < n; ++i) rd = fmax(rd, fabs(Px[i] + q[i] + Aty[i]));
        /* primal scale also counts |x| (unit-normal rows): an active row with
         * l = 0 has A x = z = 0 at the solution */
        const double sp = fmax(fmax(vmax_abs(m, Ax), vmax_abs(m, z)), vmax_abs(n, x));
        const double sd = fmax(fmax(vmax_abs(n, Px), vmax_abs(n, Aty)), vmax_abs(n, q));
        if (rp <= eps * sp && rd <= eps * sd) {
            if (qp_polish(n, rowptr, cols, P, q, m, vtx, a, l, u, z, y, xs, ys)) {
                memcpy(x, xs, sizeof(double) * n);
                memcpy(y, ys, sizeof(double) * m);
                status = ORC_OK;
                break;
            }
            if (eps > 1e-13) eps *= 1e-2;
        }
        /* adaptive rho: rho sqrt((rp / sp) / (rd / sd)), refactorised when it
         * moves by more than 5x */
        if (rp > 0 && rd > 0 && sp > 0 && sd > 0) {
            double rn = rho * sqrt((rp / sp) / (rd / sd));
            rn = rn < 1e-6 ? 1e-6 : (rn > 1e6 ? 1e6 : rn);
            if (rn > 5 * rho || rn < rho / 5) {
                rho = rn;
                for (int i = 0; i < m; ++i) rv[i] = (l[i] == u[i]) ? 1e3 * rho : rho;
                ldl_free(&f);
                fst = admm_factor(n, rowptr, cols, P, sigma, m, vtx, a, rv, work, &f);
                if (fst != ORC_OK) status = fst;
            }
        }
    }
    if (fst == ORC_OK) ldl_free(&f);
    free(work);
    free(rv);
    free(z);
    free(zp);
    free(zt);
    free(Ax);
    free(xt);
    free(rhs);
    free(Px);
    free(Aty);
    free(xs);
    free(ys);
    return status;
}

/* PlaneContact::update_active_constraints (planeContact.cpp:87-124): each
 * surface vertex takes the plane of smallest signed distance n.(x - p)
 * (plane3D.h:30-33) among those closer than PLANE_CONSTRAINT_THRESHOLD, or
 * -1. */
static void sqp_plane_active(orc_sys* s, const double* x) {
    for (int i = 0; i < s->nsurf; ++i) {
        const double* xv = x + 3 * (size_t)s->surf[i];
        double best = INFINITY;
        int idx = -1;
        for (int j = 0; j < s->nplanes; ++j) {
            double d = 0;
            for (int k = 0; k < 3; ++k) d += (xv[k] - s->plane_p[j][k]) * s->plane_n[j][k];
            if (d < ORC_PLANE_THRESHOLD && d < best) {
                best = d;
                idx = j;
            }
        }
        s->pmask[i] = idx;
    }
}

static double disk_plane_dist(const orc_sys* s, int d, const double* xv) {
    double t = 0;
    for (int k = 0; k < 3; ++k) t += s->disk_n[d][k] * (xv[k] - s->disk_c[d][k]);
    return t;
}

static double disk_radial_dist(const orc_sys* s, int d, const double* xv) {
    double v[3], t = disk_plane_dist(s, d, xv), r = 0;
    for (int k = 0; k < 3; ++k) v[k] = xv[k] - s->disk_c[d][k];
    for (int k = 0; k < 3; ++k) {
        double w = v[k] - t * s->disk_n[d][k];
        r += w * w;
    }
    return sqrt(r);
}

/* DiskContact::update_active_constraints (diskContact.cpp:91-202): a frozen
 * footprint stays until every vertex in it is at least 3 thresholds above
 * its disk's plane; otherwise, when some surface vertex is within the plane
 * threshold and the radius (+ radial threshold) of a disk, every surface
 * vertex within [-5, +1] thresholds of that disk's plane and inside its
 * radius is captured (the disk of smaller |plane distance| when two
 * compete) and the footprint freezes. */
static void sqp_disk_active(orc_sys* s, const double* x) {
    const double pt = ORC_DISK_PLANE_THRESHOLD, rt = ORC_DISK_RADIAL_THRESHOLD;
    const double clear = 3.0 * pt, depth = 5.0 * pt;
    if (s->dfrozen) {
        int all_clear = 1;
        for (int i = 0; i < s->nsurf && all_clear; ++i) {
            int d = s->dmask[i];
            if (d == -1) continue;
            if (disk_plane_dist(s, d, x + 3 * (size_t)s->surf[i]) < clear) all_clear = 0;
        }
        if (all_clear) {
            s->dfrozen = 0;
            for (int i = 0; i < s->nsurf; ++i) s->dmask[i] = -1;
        }
        return;
    }
    int hit[ORC_MAX_PRIM] = {0}, any = 0;
    for (int i = 0; i < s->nsurf; ++i) {
        const double* xv = x + 3 * (size_t)s->surf[i];
        for (int d = 0; d < s->ndisks; ++d) {
            if (disk_plane_dist(s, d, xv) >= pt) continue;
            if (disk_radial_dist(s, d, xv) > s->disk_r[d] + rt) continue;
            hit[d] = 1;
        }
    }
    for (int d = 0; d < s->ndisks; ++d) any |= hit[d];
    for (int i = 0; i < s->nsurf; ++i) s->dmask[i] = -1;
    if (!any) return;
    int nact = 0;
    for (int d = 0; d < s->ndisks; ++d) {
        if (!hit[d]) continue;
        for (int i = 0; i < s->nsurf; ++i) {
            const double* xv = x + 3 * (size_t)s->surf[i];
            double pd = disk_plane_dist(s, d, xv);
            if (pd > pt || pd < -depth) continue;
            if (disk_radial_dist(s, d, xv) > s->disk_r[d] + rt) continue;
            if (s->dmask[i] == -1) {
                s->dmask[i] = d;
                ++nact;
            } else if (fabs(pd) < fabs(disk_plane_dist(s, s->dmask[i], xv))) {
                s->dmask[i] = d;
            }
        }
    }
    if (nact > 0) s->dfrozen = 1;
}

int orc_sqp_active(orc_sys* s, const double* x, int* pmask, int* dmask, int* frozen) {
    if (!s->has_contact) return ORC_ERR_ARG;
    sqp_plane_active(s, x);
    sqp_disk_active(s, x);
    if (pmask) memcpy(pmask, s->pmask, sizeof(int) * s->nsurf);
    if (dmask) memcpy(dmask, s->dmask, sizeof(int) * s->nsurf);
    if (frozen) *frozen = s->dfrozen;
    return ORC_OK;
}

/* The stacked constraint rows of Energy::compute_constraint_bounds /
 * compute_constraint_matrix_qp (energy.cpp:394-458): NeumannBC equality
 * rows x_i - xPrev_i - dt v_i = 0 on the masked DOFs (neumannBC.cpp:4-40:
 * row e_i, l = u = -(x_i - xPrev_i - dt v_i)), then PlaneContact rows
 * n.(x - p) >= 0 (planeContact.cpp:4-84: row n, l = -n.(x - p), u = inf)
 * and DiskContact rows n.(x - c) >= 0 (diskContact.cpp:4-88) on the active
 * surface vertices.  Returns the row count; arrays sized by the caller. */
static int sqp_rows(const orc_sys* s, const double* x, double dt, int* vtx, double* a, double* l, double* u) {
    int m = 0;
    for (int i = 0; i < s->n; ++i) {
        if (!s->mask[i]) continue;
        vtx[m] = i / 3;
        for (int k = 0; k < 3; ++k) a[3 * m + k] = (k == i % 3) ? 1.0 : 0.0;
        l[m] = u[m] = -(x[i] - s->qprev[i] - dt * s->bcval[i]);
        ++m;
    }
    if (s->has_contact) {
        for (int i = 0; i < s->nsurf; ++i) {
            int j = s->pmask[i];
            if (j < 0) continue;
            const double* xv = x + 3 * (size_t)s->surf[i];
            double d = 0;
            for (int k = 0; k < 3; ++k) {
                d += s->plane_n[j][k] * (xv[k] - s->plane_p[j][k]);
                a[3 * m + k] = s->plane_n[j][k];
            }
            vtx[m] = s->surf[i];
            l[m] = -d;
            u[m] = INFINITY;
            ++m;
        }
        for (int i = 0; i < s->nsurf; ++i) {
            int j = s->dmask[i];
            if (j < 0) continue;
            const double* xv = x + 3 * (size_t)s->surf[i];
            for (int k = 0; k < 3; ++k) a[3 * m + k] = s->disk_n[j][k];
            vtx[m] = s->surf[i];
            l[m] = -disk_plane_dist(s, j, xv);
            u[m] = INFINITY;
            ++m;
        }
    }
    return m;
}

/* Solver::minimize_SQP (solver.cpp:109-197): the active sets are updated once
 * from the initial guess; then up to max_iter times the gradient and Hessian
 * (SQP assembly: no masking, no Dirichlet override) define the QP above,
 * x += dx, and ||dx|| < dx_tol ends it.  A NaN step or running out of
 * iterations is "SQP method did not converge" (ORC_ERR_NOCONV).  The QP's
 * ADMM iterations are counted in st->pcg_iters. */
static int orc_minimize_sqp(orc_sys* s, double* x, double dt, const orc_opts* o, orc_stats* st) {
    const int n = s->n;
    if (s->has_contact) {
        sqp_plane_active(s, x);
        sqp_disk_active(s, x);
    }
    orc_pattern(s, dt);
    const int mmax = n + 2 * (s->has_contact ? s->nsurf : 0);
    double* g = (double*)malloc(sizeof(double) * n);
    double* dx = (double*)malloc(sizeof(double) * n);
    double* vals = (double*)malloc(sizeof(double) * (size_t)(s->nnz > 0 ? s->nnz : 1));
    int* vtx = (int*)malloc(sizeof(int) * (mmax > 0 ? mmax : 1));
    double* a = (double*)malloc(sizeof(double) * 3 * (size_t)(mmax > 0 ? mmax : 1));
    double* l = (double*)malloc(sizeof(double) * (mmax > 0 ? mmax : 1));
    double* u = (double*)malloc(sizeof(double) * (mmax > 0 ? mmax : 1));
    double* y = (double*)malloc(sizeof(double) * (mmax > 0 ? mmax : 1));
    int status = ORC_ERR_NOCONV;
    for (int it = 0; it < o->max_iter; ++it) {
        double t0 = now_s();
        orc_gradient(s, x, dt, g);
        double t1 = now_s();
        st->t_grad += t1 - t0;
        st->gnorm = norm2(n, g);
        int hst = orc_hessian(s, x, dt, o->project, vals);
        double t2 = now_s();
        st->t_hess += t2 - t1;
        if (hst != ORC_OK) {
            status = hst;
            break;
        }
        int m = sqp_rows(s, x, dt, vtx, a, l, u);
        int qit = 0;
        int qst = orc_qp_solve(n, s->rowptr, s->cols, vals, g, m, vtx, a, l, u, dx, y, &qit);
        st->pcg_iters += qit;
        st->t_solve += now_s() - t2;
        if (qst != ORC_OK) {
            status = ORC_ERR_SOLVE;
            break;
        }
        st->newton_iters++;
        double dxn = norm2(n, dx);
        for (int i = 0; i < n; ++i) x[i] += dx[i];
        st->dxnorm = dxn;
        if (dxn != dxn) break;
        if (dxn < o->dx_tol) {
            status = ORC_OK;
            break;
        }
    }
    free(g);
    free(dx);
    free(vals);
    free(vtx);
    free(a);
    free(l);
    free(u);
    free(y);
    return status;
}

/* Solver::step (solver.cpp:5-44) without CFL substepping: dt == 0 uses the
 * stored dt given by the caller, substeps >= 1, update_state after each. */
int orc_step(orc_sys* s, double* q, double dt, int substeps, const orc_opts* o, orc_stats* st) {
    orc_opts od;
    orc_stats sd;
    if (!o) {
        orc_default_opts(&od);
        o = &od;
    }
    if (!st) st = &sd;
    memset(st, 0, sizeof(*st));
    if (substeps < 1) substeps = 1;
    int saved = s->nthreads;
    if (o->nthreads > 0) s->nthreads = o->nthreads;
    double h = dt / substeps;
    int status = ORC_OK;
    for (int j = 0; j < substeps && status == ORC_OK; ++j) {
        orc_contact_lag(s, h);
        status = orc_minimize(s, q, h, o, st);
        if (status == ORC_OK) {
            orc_update_state(s, q, h);
            for (int d = 0; d < s->ndisks; ++d)
                for (int k = 0; k < 3; ++k) s->disk_c_last[d][k] = s->disk_c[d][k];
        }
    }
    s->nthreads = saved;
    return status;
}

/* CPU-baseline sample: assemble the (projected) Hessian at q and run exactly
 * `iters` FP64 PCG iterations on it; returns seconds per PCG iteration. */
double orc_bench_pcg(orc_sys* s, const double* q, double dt, int iters, int project) {
    orc_pattern(s, dt);
    int n = s->n;
    double* vals = (double*)malloc(sizeof(double) * (size_t)(s->nnz > 0 ? s->nnz : 1));
    double* b = (double*)malloc(sizeof(double) * n);
    double* x = (double*)malloc(sizeof(double) * n);
    orc_hessian(s, q, dt, project, vals);
    for (int i = 0; i < n; ++i) b[i] = 1.0 + 1e-3 * (i % 7);
    double t0 = now_s();
    long long k = pcg_solve(n, s->rowptr, s->cols, vals, b, x, 0.0, iters, s->nthreads);
    double t = now_s() - t0;
    free(vals);
    free(b);
    free(x);
    return k > 0 ? t / (double)k : -1.0;
}

/* ---- active fibre stress (a formulation of this build; the reference's
 * actuated energies receive per-group activations the same way,
 * cpp/src/energies/softconMuscle.cpp:12-134, element.h:52) ---------------- */
int orc_set_active(orc_sys* s, int ngroups, const int* group, const double* fibre, double sigma) {
    if (s->ed != 4 || s->poly || ngroups < 1) return ORC_ERR_ARG;
    free(s->agroup);
    free(s->afib);
    free(s->act);
    s->agroup = (int*)malloc(sizeof(int) * (size_t)(s->nel > 0 ? s->nel : 1));
    s->afib = (double*)malloc(sizeof(double) * 3 * (size_t)(s->nel > 0 ? s->nel : 1));
    s->act = (double*)calloc((size_t)ngroups, sizeof(double));
    for (int e = 0; e < s->nel; ++e) {
        s->agroup[e] = group[e];
        double f[3] = {fibre[3 * (size_t)e], fibre[3 * (size_t)e + 1], fibre[3 * (size_t)e + 2]};
        double l = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
        for (int k = 0; k < 3; ++k) s->afib[3 * (size_t)e + k] = l > 0 ? f[k] / l : 0.0;
    }
    s->ngroups = ngroups;
    s->sigma = sigma;
    return ORC_OK;
}

void orc_set_activation(orc_sys* s, const double* act) {
    if (s->act) memcpy(s->act, act, sizeof(double) * (size_t)s->ngroups);
}

/* neoHookean.cpp:10-11 with the stored Poisson ratios */
void orc_set_youngs(orc_sys* s, const double* youngs) {
    for (int e = 0; e < s->nel; ++e) orc_lame(youngs[e], s->nu[e], &s->mu[e], &s->lam[e]);
}

/* ---- element energy plugins ------------------------------------------------ */
int orc_set_material(orc_sys* s, int material) {
    if (s->poly || (material != ORC_MAT_NH && material != ORC_MAT_SNH)) return ORC_ERR_ARG;
    s->material = material;
    return ORC_OK;
}

/* SoftconMuscleEnergy on the elements with group >= 0 (tetrahedron.cpp:82-88,
 * hexahedron.cpp:51-57: "muscleGroup", "softconStiffness", "softconDirection");
 * activations per group as the "softconMuscle" actuation (element.h:52) */
int orc_set_muscle(orc_sys* s, int ngroups, const int* group, const double* stiffness, const double* dir) {
    if (s->poly || ngroups < 1) return ORC_ERR_ARG;
    free(s->mgroup);
    free(s->mk);
    free(s->mdir);
    free(s->mact);
    size_t ne = (size_t)(s->nel > 0 ? s->nel : 1);
    s->mgroup = (int*)malloc(sizeof(int) * ne);
    s->mk = (double*)malloc(sizeof(double) * ne);
    s->mdir = (double*)malloc(sizeof(double) * 3 * ne);
    s->mact = (double*)calloc((size_t)ngroups, sizeof(double));
    for (int e = 0; e < s->nel; ++e) {
        s->mgroup[e] = group[e];
        s->mk[e] = stiffness[e];
        for (int k = 0; k < 3; ++k) s->mdir[3 * (size_t)e + k] = dir[3 * (size_t)e + k];
    }
    s->mngroups = ngroups;
    return ORC_OK;
}

void orc_set_muscle_activation(orc_sys* s, const double* act) {
    if (s->mact) memcpy(s->mact, act, sizeof(double) * (size_t)s->mngroups);
}
