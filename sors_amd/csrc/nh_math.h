// nh_math.h -- per-lane Neo-Hookean element math for gfx950 kernels.
//
// Structured form of the reference's element quantities.  With per-node
// shape-gradient vectors g_a (tet: rows of Dm^-1 and g_3 = -sum,
// tetrahedron.cpp:17-37; hex: G_s = hatG_s hatJ_s^-1, hexahedron.cpp:170-207)
// the reference's dense products
//     grad = V0 w dEdF * dFdx,        hess = V0 w dFdx^T K dFdx
// (neoHookean.cpp:60-70, 92-126) reduce to 3-vectors and 3x3 node blocks:
//     F      = sum_a x_a g_a^T
//     grad_a = V0 w P g_a,            P = mu F + (lam lnJ - mu) F^-T
//     H_ab   = V0 w [ mu (g_a.g_b) I + lam h_a h_b^T - c h_b h_a^T ],
//              h_a = F^-T g_a,  c = lam lnJ - mu
// which is the same tensor contraction as K = mu I + lam f f^T + c dF^-T/dF
// (d(F^-T)_ij/dF_kl = -(F^-1)_jk (F^-1)_li).
//
// SPD projection (the reference never projects; the device PCG needs it):
// the 9x9 d2Psi/dF2 is rebuilt from its analytic eigensystem (Smith, de Goes,
// Kim 2019, "Analytic Eigensystems for Isotropic Distortion Energies") with
// negative eigenvalues clamped to zero.  For compressible Neo-Hookean with
// singular values s_i of F and lnJ = ln(s1 s2 s3):
//   twist (i,j):  (dPsi_i + dPsi_j)/(s_i + s_j) = mu + (lam lnJ - mu)/(s_i s_j)
//   flip  (i,j):  (dPsi_i - dPsi_j)/(s_i - s_j) = mu + (mu - lam lnJ)/(s_i s_j)
//   scaling:      3x3 block  d2Psi/ds_i ds_j = delta_ij (mu + (mu - lam lnJ)/s_i^2)
//                                             + lam/(s_i s_j)
// Eigen-matrices: twist/flip U (e_i e_j^T -+ e_j e_i^T) V^T / sqrt(2), scaling
// U diag(z) V^T.  F is handled in displacement form F = I + Du so that the
// singular-value gaps and lnJ keep their relative accuracy in FP32.
//
// Stable Neo-Hookean (stableNeoHookean.cpp:3-150; Smith, de Goes, Kim 2018)
// with the reference's parameters mu' = 4/3 mu, lam' = lam + 5/8 mu',
// alpha = 1 + 3/4 mu'/lam' (stableNeoHookean.cpp:9-14), held per element as
// (mu', lam'), and a = alpha - 1:
//   Psi = mu'/2 (I_C - 3) + lam'/2 (J - alpha)^2 - mu'/2 log(I_C + 1)
//   P   = mu' I_C/(I_C + 1) F + lam' (J - alpha) cof F
//   K   = mu' I_C/(I_C+1) I + 2 mu'/(I_C+1)^2 f f^T + lam' (2J - alpha) J h h^T
//         + lam' (J - alpha) J dF^-T/dF          (f = vec F, h = vec F^-T)
// Its projection uses the same eigensystem as Neo-Hookean, with
// dPsi/dI_C = mu'/2 I_C/(I_C+1) and dPsi/dJ = lam' (J - alpha), and signed
// singular values (an inverted element has s_min < 0; the energy is finite).
//
// SoftCon muscle (softconMuscle.cpp:12-134) per Gauss point, with the
// element's direction m, stiffness k, group activation e, u = F m, l = |u|,
// L = 1 - e, r = L / l:
//   Psi = k/2 |l - L|                          (the reference's energy, :31)
//   P   = k (1 - r) u m^T                      (its gradient, :55-72: that of
//                                               k/2 (l - L)^2, not of Psi)
//   K   = k [ (r/l^2) u u^T + (1 - r) 1 1^T ] (x) m m^T
//                                              (its Hessian, :98-130, with
//                                               1 = (1,1,1): the second term
//                                               sums over all rows of dF/dx)
// The projected K clamps the 3x3 factor M = (r/l^2) u u^T + (1 - r) 1 1^T:
// the eigenvalues of M (x) m m^T are |m|^2 times those of M.
#pragma once

#include <hip/hip_runtime.h>

#define SORS_HD __host__ __device__ __forceinline__

namespace sors {

template <typename T>
struct M3 {
    T m[9];
    SORS_HD T& operator()(int i, int j) { return m[3 * i + j]; }
    SORS_HD const T& operator()(int i, int j) const { return m[3 * i + j]; }
};

template <typename T>
SORS_HD T det3(const M3<T>& A) {
    return A.m[0] * (A.m[4] * A.m[8] - A.m[5] * A.m[7]) - A.m[1] * (A.m[3] * A.m[8] - A.m[5] * A.m[6]) +
           A.m[2] * (A.m[3] * A.m[7] - A.m[4] * A.m[6]);
}

// cofactor matrix C with C_ij = d(det)/dA_ij, so A^-T = C / det
template <typename T>
SORS_HD M3<T> cofactor3(const M3<T>& A) {
    M3<T> C;
    C.m[0] = A.m[4] * A.m[8] - A.m[5] * A.m[7];
    C.m[1] = A.m[5] * A.m[6] - A.m[3] * A.m[8];
    C.m[2] = A.m[3] * A.m[7] - A.m[4] * A.m[6];
    C.m[3] = A.m[2] * A.m[7] - A.m[1] * A.m[8];
    C.m[4] = A.m[0] * A.m[8] - A.m[2] * A.m[6];
    C.m[5] = A.m[1] * A.m[6] - A.m[0] * A.m[7];
    C.m[6] = A.m[1] * A.m[5] - A.m[2] * A.m[4];
    C.m[7] = A.m[2] * A.m[3] - A.m[0] * A.m[5];
    C.m[8] = A.m[0] * A.m[4] - A.m[1] * A.m[3];
    return C;
}

// J - 1 for F = I + D, without forming F: tr D + (tr^2 D - tr D^2)/2 + det D
template <typename T>
SORS_HD T jm1_from_disp(const M3<T>& D) {
    T tr = D.m[0] + D.m[4] + D.m[8];
    T i2 = D.m[0] * D.m[4] + D.m[4] * D.m[8] + D.m[0] * D.m[8] - D.m[1] * D.m[3] - D.m[2] * D.m[6] -
           D.m[5] * D.m[7];
    return tr + i2 + det3(D);
}

SORS_HD double log1p_t(double x) { return log1p(x); }
SORS_HD float log1p_t(float x) { return log1pf(x); }
SORS_HD double sqrt_t(double x) { return sqrt(x); }
SORS_HD float sqrt_t(float x) { return sqrtf(x); }
SORS_HD double abs_t(double x) { return fabs(x); }
SORS_HD float abs_t(float x) { return fabsf(x); }

// Psi of compressible Neo-Hookean (neoHookean.cpp:31-37) in displacement
// form: tr(F^T F) - 3 = 2 tr D + |D|^2, lnJ = log1p(J - 1).
template <typename T>
SORS_HD T nh_psi_disp(const M3<T>& D, T mu, T lam, bool* inverted) {
    T jm1 = jm1_from_disp(D);
    if (!(jm1 > T(-1))) *inverted = true;
    T lj = log1p_t(jm1);
    T n2 = 0;
    for (int i = 0; i < 9; ++i) n2 += D.m[i] * D.m[i];
    T tr = D.m[0] + D.m[4] + D.m[8];
    return T(0.5) * mu * (T(2) * tr + n2) - mu * lj + T(0.5) * lam * lj * lj;
}

// Per-Gauss-point stress data shared by gradient and Hessian.
template <typename T>
struct NHPoint {
    M3<T> F;     // deformation gradient
    M3<T> FiT;   // F^-T
    T lnJ;
    T J;
};

template <typename T>
SORS_HD NHPoint<T> nh_point(const M3<T>& D) {
    NHPoint<T> p;
    for (int i = 0; i < 9; ++i) p.F.m[i] = D.m[i];
    p.F.m[0] += T(1);
    p.F.m[4] += T(1);
    p.F.m[8] += T(1);
    T jm1 = jm1_from_disp(D);
    p.J = T(1) + jm1;
    p.lnJ = log1p_t(jm1);
    M3<T> C = cofactor3(p.F);
    T iJ = T(1) / p.J;
    for (int i = 0; i < 9; ++i) p.FiT.m[i] = C.m[i] * iJ;
    return p;
}

// first Piola-Kirchhoff stress (neoHookean.cpp:68)
template <typename T>
SORS_HD M3<T> nh_stress(const NHPoint<T>& p, T mu, T lam) {
    M3<T> P;
    T c = lam * p.lnJ - mu;
    for (int i = 0; i < 9; ++i) P.m[i] = mu * p.F.m[i] + c * p.FiT.m[i];
    return P;
}

// ---------------------------------------------------------------------------
// Symmetric 3x3 eigen-decomposition by cyclic Jacobi (fixed sweep count).
// A is overwritten with its diagonal form; V collects the rotations (columns
// are eigenvectors).
template <typename T>
SORS_HD void jacobi3(T A[3][3], T V[3][3]) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) V[i][j] = (i == j) ? T(1) : T(0);
#pragma unroll
    for (int sweep = 0; sweep < 6; ++sweep) {
#pragma unroll
        for (int pq = 0; pq < 3; ++pq) {
            const int p = (pq == 2) ? 1 : 0;
            const int q = (pq == 0) ? 1 : 2;
            T apq = A[p][q];
            if (apq == T(0)) continue;
            T theta = (A[q][q] - A[p][p]) / (T(2) * apq);
            T t = T(1) / (abs_t(theta) + sqrt_t(theta * theta + T(1)));
            if (theta < T(0)) t = -t;
            T c = T(1) / sqrt_t(t * t + T(1));
            T s = t * c;
            for (int k = 0; k < 3; ++k) {
                T akp = A[k][p], akq = A[k][q];
                A[k][p] = c * akp - s * akq;
                A[k][q] = s * akp + c * akq;
            }
            for (int k = 0; k < 3; ++k) {
                T apk = A[p][k], aqk = A[q][k];
                A[p][k] = c * apk - s * aqk;
                A[q][k] = s * apk + c * aqk;
            }
            for (int k = 0; k < 3; ++k) {
                T vkp = V[k][p], vkq = V[k][q];
                V[k][p] = c * vkp - s * vkq;
                V[k][q] = s * vkp + c * vkq;
            }
        }
    }
}

// Projected (PSD) d2Psi/dF2 of Neo-Hookean as a packed symmetric 9x9
// (row-major upper triangle, 45 entries, index (3i+j, 3k+l)).
template <typename T>
struct Sym9 {
    T a[45];
    SORS_HD static int idx(int r, int c) {
        if (r > c) {
            int t = r;
            r = c;
            c = t;
        }
        return r * 9 - (r * (r - 1)) / 2 + (c - r);
    }
    SORS_HD T get(int r, int c) const { return a[idx(r, c)]; }
};

template <typename T>
SORS_HD void sym9_add_outer(Sym9<T>& S, const T q[9], T lam) {
    int k = 0;
    for (int r = 0; r < 9; ++r)
        for (int c = r; c < 9; ++c) S.a[k++] += lam * q[r] * q[c];
}

// Analytic eigensystem of Neo-Hookean d2Psi/dF2 at F = I + D, clamped.
template <typename T>
SORS_HD void nh_projected_tangent(const M3<T>& D, const NHPoint<T>& p, T mu, T lam, Sym9<T>& S) {
    for (int i = 0; i < 45; ++i) S.a[i] = T(0);
    // E = F^T F - I = D + D^T + D^T D, eigen-decomposed: E = V diag(e) V^T
    T A[3][3], V[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            T acc = D(i, j) + D(j, i);
            for (int k = 0; k < 3; ++k) acc += D(k, i) * D(k, j);
            A[i][j] = acc;
        }
    jacobi3(A, V);
    T e[3] = {A[0][0], A[1][1], A[2][2]};
    T s[3];
    for (int i = 0; i < 3; ++i) s[i] = sqrt_t(T(1) + e[i]);
    // U = F V S^-1 (F = U S V^T with det U = det V * sign(det F) > 0 path)
    T U[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            T acc = 0;
            for (int k = 0; k < 3; ++k) acc += p.F(i, k) * V[k][j];
            U[i][j] = acc / s[j];
        }
    const T lj = p.lnJ;
    // twist / flip pairs
    const T r2 = T(0.70710678118654752440);
    for (int pair = 0; pair < 3; ++pair) {
        const int i = (pair == 2) ? 1 : 0;
        const int j = (pair == 0) ? 1 : 2;
        T ss = s[i] * s[j];
        // s_i s_j - 1 = (e_i + e_j + e_i e_j) / (s_i s_j + 1), no cancellation
        T ssm1 = (e[i] + e[j] + e[i] * e[j]) / (ss + T(1));
        T lt = (mu * ssm1 + lam * lj) / ss;            // twist
        T lf = (mu * (ss + T(1)) - lam * lj) / ss;     // flip
        if (lt > T(0) || lf > T(0)) {
            T qt[9], qf[9];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    T ui = U[a][i], uj = U[a][j];
                    T vi = V[b][i], vj = V[b][j];
                    qt[3 * a + b] = r2 * (ui * vj - uj * vi);
                    qf[3 * a + b] = r2 * (ui * vj + uj * vi);
                }
            if (lt > T(0)) sym9_add_outer(S, qt, lt);
            if (lf > T(0)) sym9_add_outer(S, qf, lf);
        }
    }
    // scaling block
    T B[3][3], Z[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            B[i][j] = lam / (s[i] * s[j]);
            if (i == j) B[i][j] += mu + (mu - lam * lj) / (s[i] * s[i]);
        }
    jacobi3(B, Z);
    for (int k = 0; k < 3; ++k) {
        T lk = B[k][k];
        if (!(lk > T(0))) continue;
        T q[9];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                T acc = 0;
                for (int m = 0; m < 3; ++m) acc += U[a][m] * Z[m][k] * V[b][m];
                q[3 * a + b] = acc;
            }
        sym9_add_outer(S, q, lk);
    }
}

// H_ab block (3x3, row-major) from the projected tangent:
// (H_ab)_ik = sum_jl ga_j S(3i+j, 3k+l) gb_l
template <typename T>
SORS_HD void block_from_tangent(const Sym9<T>& S, const T ga[3], const T gb[3], T w, T out[9]) {
    for (int i = 0; i < 3; ++i) {
        T t[9];  // t[(k,l)] = sum_j ga_j S(3i+j, 3k+l)
        for (int kl = 0; kl < 9; ++kl) {
            T acc = 0;
            for (int j = 0; j < 3; ++j) acc += ga[j] * S.get(3 * i + j, kl);
            t[kl] = acc;
        }
        for (int k = 0; k < 3; ++k) {
            T acc = 0;
            for (int l = 0; l < 3; ++l) acc += t[3 * k + l] * gb[l];
            out[3 * i + k] += w * acc;
        }
    }
}

// H_ab block of the raw (unprojected) reference tangent.
template <typename T>
SORS_HD void block_raw(T mu, T lam, T c, const T ga[3], const T gb[3], const T ha[3], const T hb[3], T w,
                       T out[9]) {
    T gg = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
            T v = lam * ha[i] * hb[k] - c * ha[k] * hb[i];
            if (i == k) v += mu * gg;
            out[3 * i + k] += w * v;
        }
}


// ---------------------------------------------------------------------------
// Raw 9x9 tangent of an isotropic energy as a packed symmetric Sym9:
//   K(3i+j, 3p+q) = c0 d_ip d_jq + c1 F_ij F_pq + c2 h_ij h_pq - c3 h_iq h_pj
// (h = F^-T; -h_iq h_pj is d(F^-T)_ij / dF_pq, neoHookean.cpp:104-110).
// Neo-Hookean: c0 = mu, c1 = 0, c2 = lam, c3 = lam lnJ - mu.
template <typename T>
SORS_HD void raw_tangent_sym9(T c0, T c1, T c2, T c3, const M3<T>& F, const M3<T>& h, Sym9<T>& S) {
    int k = 0;
    for (int r = 0; r < 9; ++r) {
        const int i = r / 3, j = r % 3;
        for (int c = r; c < 9; ++c) {
            const int p = c / 3, q = c % 3;
            T v = c1 * F.m[r] * F.m[c] + c2 * h.m[r] * h.m[c] - c3 * h(i, q) * h(p, j);
            if (r == c) v += c0;
            S.a[k++] = v;
        }
    }
}

// ---- Stable Neo-Hookean ----------------------------------------------------
SORS_HD double log_t(double x) { return log(x); }
SORS_HD float log_t(float x) { return logf(x); }

// a = alpha - 1 = 3/4 mu'/lam'
template <typename T>
SORS_HD T snh_a(T mu, T lam) {
    return T(0.75) * mu / lam;
}

// Psi(F) - Psi(I) in displacement form F = I + D: I_C - 3 = 2 tr D + |D|^2,
// (J - alpha)^2 - a^2 = (J - 1)(J - 1 - 2a), log(I_C + 1) - log 4 = log1p((I_C - 3)/4)
template <typename T>
SORS_HD T snh_psi_shift(const M3<T>& D, T mu, T lam) {
    const T jm1 = jm1_from_disp(D);
    T n2 = 0;
    for (int i = 0; i < 9; ++i) n2 += D.m[i] * D.m[i];
    const T ic3 = T(2) * (D.m[0] + D.m[4] + D.m[8]) + n2;
    const T a = snh_a(mu, lam);
    return T(0.5) * mu * ic3 + T(0.5) * lam * jm1 * (jm1 - T(2) * a) - T(0.5) * mu * log1p_t(ic3 * T(0.25));
}

// Psi(I) = lam'/2 a^2 - mu'/2 log 4 (the reference does not subtract it)
template <typename T>
SORS_HD T snh_psi_rest(T mu, T lam) {
    const T a = snh_a(mu, lam);
    return T(0.5) * lam * a * a - T(0.5) * mu * T(1.38629436111989061883);
}

// first Piola stress (stableNeoHookean.cpp:64-77): mu' I_C/(I_C+1) F + lam' (J - alpha) cof F
template <typename T>
SORS_HD M3<T> snh_stress(const NHPoint<T>& p, const M3<T>& D, T mu, T lam) {
    T n2 = 0;
    for (int i = 0; i < 9; ++i) n2 += D.m[i] * D.m[i];
    const T ic = T(3) + T(2) * (D.m[0] + D.m[4] + D.m[8]) + n2;
    const T c0 = mu * ic / (ic + T(1));
    const T cj = lam * ((p.J - T(1)) - snh_a(mu, lam));   // lam' (J - alpha)
    const M3<T> C = cofactor3(p.F);
    M3<T> P;
    for (int i = 0; i < 9; ++i) P.m[i] = c0 * p.F.m[i] + cj * C.m[i];
    return P;
}

// raw 9x9 tangent (stableNeoHookean.cpp:112-131)
template <typename T>
SORS_HD void snh_raw_tangent(const NHPoint<T>& p, const M3<T>& D, T mu, T lam, Sym9<T>& S) {
    T n2 = 0;
    for (int i = 0; i < 9; ++i) n2 += D.m[i] * D.m[i];
    const T ic = T(3) + T(2) * (D.m[0] + D.m[4] + D.m[8]) + n2;
    const T jma = (p.J - T(1)) - snh_a(mu, lam);  // J - alpha
    const T c0 = mu * ic / (ic + T(1));
    const T c1 = T(2) * mu / ((ic + T(1)) * (ic + T(1)));
    const T c2 = lam * (jma + p.J) * p.J;           // lam' (2J - alpha) J
    const T c3 = lam * jma * p.J;                   // lam' (J - alpha) J
    raw_tangent_sym9(c0, c1, c2, c3, p.F, p.FiT, S);
}

// Projected tangent from the analytic eigensystem, singular values signed so
// that U and V are rotations.  Per pair (i, j) with k the third index:
//   twist  2 Psi_2 + Psi_3 s_k = mu' I_C/(I_C+1) + lam'(J - alpha) s_k
//   flip   2 Psi_2 - Psi_3 s_k
//   scaling d2Psi/ds_i ds_j = d_ij mu' I_C/(I_C+1) + 2 mu' s_i s_j/(I_C+1)^2
//                             + lam' (J/s_i)(J/s_j) + (1 - d_ij) lam'(J - alpha) s_k
// (Psi_2 = dPsi/dI_C, Psi_3 = dPsi/dJ).  The twist value is written without
// cancellation near rest: mu' (I_C - 3)/(4 (I_C + 1)) - 3/4 mu' (s_k - 1)
// + lam' (J - 1) s_k.
template <typename T>
SORS_HD void snh_projected_tangent(const M3<T>& D, const NHPoint<T>& p, T mu, T lam, Sym9<T>& S) {
    for (int i = 0; i < 45; ++i) S.a[i] = T(0);
    T A[3][3], V[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            T acc = D(i, j) + D(j, i);
            for (int k = 0; k < 3; ++k) acc += D(k, i) * D(k, j);
            A[i][j] = acc;
        }
    jacobi3(A, V);
    T e[3] = {A[0][0], A[1][1], A[2][2]};
    T s[3];
    for (int i = 0; i < 3; ++i) s[i] = sqrt_t(T(1) + e[i]);
    T U[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            T acc = 0;
            for (int k = 0; k < 3; ++k) acc += p.F(i, k) * V[k][j];
            U[i][j] = acc / s[j];
        }
    const T jm1 = jm1_from_disp(D);
    if (jm1 < T(-1)) {  // inverted: the smallest singular value takes the sign of J
        int km = 0;
        for (int k = 1; k < 3; ++k)
            if (s[k] < s[km]) km = k;
        s[km] = -s[km];
        for (int i = 0; i < 3; ++i) U[i][km] = -U[i][km];
    }
    const T J = T(1) + jm1;
    const T a = snh_a(mu, lam);
    const T ic3 = e[0] + e[1] + e[2];
    const T ic1 = ic3 + T(4);                  // I_C + 1
    const T c0 = mu * (ic3 + T(3)) / ic1;      // mu' I_C/(I_C+1)
    const T r2 = T(0.70710678118654752440);
    for (int pair = 0; pair < 3; ++pair) {
        const int i = (pair == 2) ? 1 : 0;
        const int j = (pair == 0) ? 1 : 2;
        const int k = 3 - i - j;
        const T skm1 = s[k] > T(0) ? e[k] / (s[k] + T(1)) : s[k] - T(1);
        const T lt = mu * ic3 / (T(4) * ic1) - T(0.75) * mu * skm1 + lam * jm1 * s[k];  // twist
        const T lf = c0 - lam * (jm1 - a) * s[k];                                       // flip
        if (lt > T(0) || lf > T(0)) {
            T qt[9], qf[9];
            for (int aa = 0; aa < 3; ++aa)
                for (int b = 0; b < 3; ++b) {
                    T ui = U[aa][i], uj = U[aa][j];
                    T vi = V[b][i], vj = V[b][j];
                    qt[3 * aa + b] = r2 * (ui * vj - uj * vi);
                    qf[3 * aa + b] = r2 * (ui * vj + uj * vi);
                }
            if (lt > T(0)) sym9_add_outer(S, qt, lt);
            if (lf > T(0)) sym9_add_outer(S, qf, lf);
        }
    }
    T B[3][3], Z[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const T Ji = s[(i + 1) % 3] * s[(i + 2) % 3], Jj = s[(j + 1) % 3] * s[(j + 2) % 3];
            T v = T(2) * mu * s[i] * s[j] / (ic1 * ic1) + lam * Ji * Jj;
            if (i == j)
                v += c0;
            else
                v += lam * (J - T(1) - a) * s[3 - i - j];
            B[i][j] = v;
        }
    jacobi3(B, Z);
    for (int k = 0; k < 3; ++k) {
        T lk = B[k][k];
        if (!(lk > T(0))) continue;
        T q[9];
        for (int aa = 0; aa < 3; ++aa)
            for (int b = 0; b < 3; ++b) {
                T acc = 0;
                for (int m = 0; m < 3; ++m) acc += U[aa][m] * Z[m][k] * V[b][m];
                q[3 * aa + b] = acc;
            }
        sym9_add_outer(S, q, lk);
    }
}

// ---- SoftCon muscle ----------------------------------------------------------
// u = F m with F = I + D
template <typename T>
SORS_HD void muscle_u(const M3<T>& D, const T m[3], T u[3]) {
    for (int i = 0; i < 3; ++i) u[i] = m[i] + D(i, 0) * m[0] + D(i, 1) * m[1] + D(i, 2) * m[2];
}

// Psi = k/2 |l - L| (softconMuscle.cpp:27-31)
template <typename T>
SORS_HD T muscle_psi(const M3<T>& D, const T m[3], T k, T L) {
    T u[3];
    muscle_u(D, m, u);
    const T l = sqrt_t(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    return T(0.5) * k * abs_t(l - L);
}

// P += k (1 - r) u m^T (softconMuscle.cpp:55-72)
template <typename T>
SORS_HD void muscle_add_stress(const M3<T>& D, const T m[3], T k, T L, M3<T>& P) {
    T u[3];
    muscle_u(D, m, u);
    const T l = sqrt_t(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const T c = k * (T(1) - L / l);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) P(i, j) += c * u[i] * m[j];
}

// S += k M (x) m m^T, M = (r/l^2) u u^T + (1 - r) 1 1^T, clamped to PSD when
// project (softconMuscle.cpp:98-130)
template <typename T>
SORS_HD void muscle_add_tangent(const M3<T>& D, const T m[3], T k, T L, bool project, Sym9<T>& S) {
    T u[3];
    muscle_u(D, m, u);
    const T l2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const T l = sqrt_t(l2);
    const T r = L / l;
    T M[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M[i][j] = r / l2 * u[i] * u[j] + (T(1) - r);
    if (project) {
        T A[3][3], Z[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) A[i][j] = M[i][j];
        jacobi3(A, Z);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                T acc = 0;
                for (int q = 0; q < 3; ++q)
                    if (A[q][q] > T(0)) acc += A[q][q] * Z[i][q] * Z[j][q];
                M[i][j] = acc;
            }
    }
    int idx = 0;
    for (int rr = 0; rr < 9; ++rr) {
        const int i = rr / 3, j = rr % 3;
        for (int c = rr; c < 9; ++c) {
            const int p = c / 3, q = c % 3;
            S.a[idx++] += k * M[i][p] * m[j] * m[q];
        }
    }
}

}  // namespace sors
