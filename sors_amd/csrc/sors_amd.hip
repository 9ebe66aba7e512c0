// sors_amd.hip -- MI355X implicit Newton step: element kernels, fixed-CSR
// assembly, block-Jacobi PCG, device line search, and the C ABI.
//
// Reference path restated here (cpp/... paths of srl-ethz/sors):
//   Solver::step / minimize_newton / minimize_newton_linesearch
//       (cpp/src/solver.cpp:5-44, 56-105, 201-264)
//   Energy::compute_energy / compute_gradient / compute_hessian
//       (cpp/src/energy.cpp:208-248, 252-308, 312-390)
//   TimeIntegrator::integrate_* / update_state
//       (cpp/src/integrator.cpp:3-124, cpp/headers/integrator.h:79-97)
//   Tetrahedron / Hexahedron / NeoHookeanEnergy element math
//       (cpp/src/geometries/tetrahedron.cpp, hexahedron.cpp, cpp/src/energies/neoHookean.cpp)
// The direct SimplicialLDLT solve (solver.cpp:82-84) is replaced by a
// block-Jacobi preconditioned conjugate gradient on the fixed CSR pattern,
// with SPD-projected element tangents.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "coarse.h"
#include "kernels.h"
#include "nh_math.h"
#include "pattern.h"
#include "sors_amd.h"

namespace sors {

// ===========================================================================
// Reductions: deterministic two-level sums.  A block reduces its values with
// a fixed tree (wavefront shuffles, then the wave sums in order) and writes
// one partial per block; the consumer (the next kernel's prologue, or a
// one-block finalize) sums the partials in a fixed order.
// ===========================================================================
constexpr int kMaxWaves = 1024 / 64;  // largest workgroup on gfx950
// fixed-order sum over the workgroup; the totals are left in thread 0 only
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
    __shared__ double sh[NV][kMaxWaves];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double x = v[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        v[k] = x;
    }
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) sh[k][wid] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = 0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[k][w];
            v[k] = s;
        }
    __syncthreads();
}

// Each block writes its NV sums to partials[k * kMaxRedBlocks + blockIdx.x]
// with plain stores; the partials become visible to the next kernel on the
// stream (kernel boundaries publish them across XCDs).  No fences: an
// agent-scope release on gfx950 writes back the XCD's whole L2.
template <int NV>
__device__ __forceinline__ void block_partial(double (&v)[NV], double* partials) {
    block_sum<NV>(v);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) partials[k * kMaxRedBlocks + blockIdx.x] = v[k];
}

// Sum of the first nb partial slots, identical (same order) in every block
// and valid in every thread.
template <int NV>
__device__ __forceinline__ void sum_partials(const double* partials, int nb, double (&out)[NV]) {
    __shared__ double bc[NV];
    double s[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) s[k] = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
#pragma unroll
        for (int k = 0; k < NV; ++k) s[k] += partials[k * kMaxRedBlocks + i];
    block_sum<NV>(s);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) bc[k] = s[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = bc[k];
    __syncthreads();
}

// one block: out[k] = sum of the nb partials of slot k
template <int NV>
__global__ __launch_bounds__(kBlock) void k_finalize(const double* partials, int nb, double* out) {
    double t[NV];
    sum_partials<NV>(partials, nb, t);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) out[k] = t[k];
}

static inline int red_grid(long long work, int per_block) {
    long long b = (work + per_block - 1) / per_block;
    if (b < 1) b = 1;
    return (int)std::min<long long>(b, kMaxRedBlocks);
}

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of logical blocks so that neighbouring rows / elements
// share that XCD's L2.
__device__ __forceinline__ int xcd_block() {
    const int nb = gridDim.x, b = blockIdx.x;
    const int per = (nb + kNumXcd - 1) / kNumXcd;
    const int xcd = b % kNumXcd, local = b / kNumXcd;
    const int full = nb % kNumXcd;  // when nb is not a multiple of 8 the tail XCDs get one block fewer
    int logical;
    if (full == 0 || xcd < full)
        logical = xcd * per + local;
    else
        logical = full * per + (xcd - full) * (per - 1) + local;
    return logical;
}

// ===========================================================================
// Fill / vector kernels
// ===========================================================================
template <typename T>
__global__ void k_fill(T* p, long long n, T v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

template <typename T>
static void fill(T* p, long long n, T v, hipStream_t st) {
    if (n <= 0) return;
    int grid = (int)std::min<long long>((n + kBlock - 1) / kBlock, 8192);
    k_fill<T><<<grid, kBlock, 0, st>>>(p, n, v);
}

// ===========================================================================
// Element kernels
// ===========================================================================
template <typename HT>
__device__ __forceinline__ void atomic_add_ht(HT* p, HT v) {
    atomicAdd(p, v);
}

// gather node positions: x = q + alpha*dq, X = rest
__device__ __forceinline__ void gather_node(const ElemArgs& a, int node, double X[3], double x[3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        X[c] = a.rest[3 * node + c];
        double v = a.q[3 * node + c];
        if (a.dq) v += a.alpha * a.dq[3 * node + c];
        x[c] = v;
    }
}

// Polynomial test element (cpp/src/_test.cpp:8-66): energy sum x^2 (x^2 - 1/2)
// over the element coordinates, not scaled by volume.
__device__ __forceinline__ double poly_psi(double x) { return x * x * (x * x - 0.5); }
__device__ __forceinline__ double poly_dpsi(double x) { return (2 * x - 1) * (2 * x + 1) * x; }
__device__ __forceinline__ double poly_d2psi(double x) { return 12 * x * x - 1; }

// scatter one 3x3 node block into the CSR values
template <typename HT>
__device__ __forceinline__ void scatter_block(const ElemArgs& a, int e, int la, int lb, int na, const HT blk[9]) {
    HT* vals = (HT*)a.vals;
    if (a.sbase) {
        const int rk = a.rank[(size_t)(la * a.ed + lb) * a.E + e];
        if (rk == kRankSkip) return;
        const long long base = block_index(a.sbase, na, rk, 0, 0);
#pragma unroll
        for (int k = 0; k < 9; ++k) atomic_add_ht(vals + base + (long long)kSlice * k, blk[k]);
    } else if (a.rank) {
        const int rk = a.rank[(size_t)(la * a.ed + lb) * a.E + e];
        if (rk == kRankSkip) return;
        const long long base = a.rowptr[3 * na];
        const long long len = a.rowptr[3 * na + 1] - base;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d) atomic_add_ht(vals + base + c * len + 3 * rk + d, blk[3 * c + d]);
    } else {
        const int nd = 3 * a.ed;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const int idx = a.fullmap[((size_t)e * nd + 3 * la + c) * nd + 3 * lb + d];
                if (idx >= 0) atomic_add_ht(vals + idx, blk[3 * c + d]);
            }
    }
}

__device__ __forceinline__ bool dof_kept(const ElemArgs& a, int dof) { return a.sqp || !a.mask || a.mask[dof] == 0; }

// active fibre stress of element e: returns sigma_max * a (0 if passive),
// the fibre direction in f
__device__ __forceinline__ double active_level(const ElemArgs& a, int e, double f[3]) {
    if (!a.agroup) return 0.0;
    const int gi = a.agroup[e];
    if (gi < 0) return 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) f[c] = a.afib[(size_t)c * a.E + e];
    return a.sigma * a.act[gi];
}

// ---- element energy plugins: Stable Neo-Hookean and SoftCon muscle -------
// SoftCon muscle of element e: returns its stiffness k (0: none), its
// direction in m and the target length L = 1 - activation
// (softconMuscle.cpp:27-30)
__device__ __forceinline__ double muscle_of(const ElemArgs& a, int e, double m[3], double* L) {
    if (!a.mgroup) return 0.0;
    const int gi = a.mgroup[e];
    if (gi < 0) return 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) m[c] = a.mdir[(size_t)c * a.E + e];
    *L = 1.0 - a.mact[gi];
    return a.mk[e];
}

__device__ __forceinline__ bool snh_material(const ElemArgs& a) { return a.material == SORS_STABLE_NEOHOOKEAN; }

// elastic first Piola stress (FP64) of the context's material, plus the muscle
__device__ __forceinline__ M3<double> elem_stress(const ElemArgs& a, const NHPoint<double>& p, const M3<double>& D,
                                                  double mu, double lam, double mk, const double m[3], double mL) {
    M3<double> P = snh_material(a) ? snh_stress(p, D, mu, lam) : nh_stress(p, mu, lam);
    if (mk != 0.0) muscle_add_stress(D, m, mk, mL, P);
    return P;
}

// energy density (FP64) of the material plus the muscle; Stable Neo-Hookean's
// rest value Psi(I) only when absolute (energy differences leave it out, so
// that it does not cancel in rounding)
__device__ __forceinline__ double elem_psi(const ElemArgs& a, const M3<double>& D, double mu, double lam, bool* inv,
                                           bool absolute, double mk, const double m[3], double mL) {
    double en;
    if (snh_material(a)) {
        en = snh_psi_shift(D, mu, lam);
        if (absolute) en += snh_psi_rest(mu, lam);
    } else {
        en = nh_psi_disp(D, mu, lam, inv);
    }
    if (mk != 0.0) en += muscle_psi(D, m, mk, mL);
    return en;
}

// 9x9 tangent (HT) of the material, projected (PROJ) or raw, plus the muscle's
// (projected on its own when PROJ), as a packed Sym9
template <typename HT, int PROJ>
__device__ __forceinline__ void elem_tangent(const ElemArgs& a, const M3<HT>& D, HT mu, HT lam, double mk,
                                             const double m[3], double mL, Sym9<HT>& S) {
    const NHPoint<HT> p = nh_point(D);
    if (snh_material(a)) {
        if (PROJ)
            snh_projected_tangent(D, p, mu, lam, S);
        else
            snh_raw_tangent(p, D, mu, lam, S);
    } else {
        if (PROJ)
            nh_projected_tangent(D, p, mu, lam, S);
        else
            raw_tangent_sym9(mu, HT(0), lam, lam * p.lnJ - mu, p.F, p.FiT, S);
    }
    if (mk != 0.0) {
        const HT mh[3] = {(HT)m[0], (HT)m[1], (HT)m[2]};
        muscle_add_tangent(D, mh, (HT)mk, (HT)mL, PROJ != 0, S);
    }
}

// ---- tet4: one element per lane ------------------------------------------
// MODE 0: assemble gradient + Hessian; MODE 1: gradient only; MODE 2: export
// per-element energy / gradient / Hessian (FP64 gradient & energy; Hessian in HT)
template <typename HT, int PROJ, int MODE>
__global__ __launch_bounds__(kBlock) void k_tet(ElemArgs a) {
    const int e = xcd_block() * blockDim.x + threadIdx.x;
    if (e >= a.E) return;
    int nd[4];
    double X[4][3], x[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        nd[k] = a.conn[(size_t)k * a.E + e];
        gather_node(a, nd[k], X[k], x[k]);
    }
    if (a.material == SORS_TEST_POLY) {
        double en = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double v = x[k][c];
                en += poly_psi(v);
                const double gv = poly_dpsi(v);
                if (MODE == 2) {
                    if (a.ex_grad) a.ex_grad[(size_t)e * 12 + 3 * k + c] = gv;
                } else if (a.gint && dof_kept(a, 3 * nd[k] + c)) {
                    atomicAdd(a.gint + 3 * nd[k] + c, gv);
                }
            }
        if (MODE == 2 && a.ex_energy) a.ex_energy[e] = en;
        if (MODE == 2 && a.ex_hess) {
            for (int i = 0; i < 144; ++i) a.ex_hess[(size_t)e * 144 + i] = 0;
            for (int i = 0; i < 12; ++i) a.ex_hess[(size_t)e * 144 + 13 * i] = poly_d2psi(x[i / 3][i % 3]);
        }
        if (MODE == 0 && a.vals) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                HT blk[9];
#pragma unroll
                for (int i = 0; i < 9; ++i) blk[i] = 0;
#pragma unroll
                for (int c = 0; c < 3; ++c) blk[4 * c] = (HT)(a.hscale * poly_d2psi(x[k][c]));
                // the diagonal block (k,k) carries everything; zero blocks elsewhere keep the pattern
                scatter_block<HT>(a, e, k, k, nd[k], blk);
            }
        }
        return;
    }
    // Dm = [X0-X3, X1-X3, X2-X3] (tetrahedron.cpp:11-15), Dm^-1 by cofactors
    M3<double> Dm;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) Dm(r, c) = X[c][r] - X[3][r];
    const M3<double> Cm = cofactor3(Dm);
    const double detm = Dm.m[0] * Cm.m[0] + Dm.m[1] * Cm.m[1] + Dm.m[2] * Cm.m[2];
    const double V0 = fabs(detm) / 6.0;
    // g_a = row a of Dm^-1 = column a of cof(Dm) / det; g_3 = -(g_0+g_1+g_2)
    double g[4][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) g[k][c] = Cm(c, k) / detm;
#pragma unroll
    for (int c = 0; c < 3; ++c) g[3][c] = -(g[0][c] + g[1][c] + g[2][c]);
    // displacement gradient D = sum_a (x_a - X_a) g_a^T, F = I + D
    M3<double> D;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double acc = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) acc += (x[k][r] - X[k][r]) * g[k][c];
            D(r, c) = acc;
        }
    const double mu = a.mu[e], lam = a.lam[e];
    const NHPoint<double> p = nh_point(D);
    if (!(p.J > 0) && !snh_material(a)) atomicOr(a.flag, 1);
    // SoftCon muscle; gen: material or muscle beyond plain Neo-Hookean
    double mdir[3] = {0.0, 0.0, 0.0}, mL = 0.0;
    const double mk = muscle_of(a, e, mdir, &mL);
    const bool gen = snh_material(a) || mk != 0.0;
    // active fibre stress: P += sigma a (F f) f^T, H_ab += V0 sigma a (f.g_a)(f.g_b) I
    double fa[3] = {0.0, 0.0, 0.0}, Ff[3] = {0.0, 0.0, 0.0}, fg[4] = {0.0, 0.0, 0.0, 0.0};
    const double sa = active_level(a, e, fa);
    if (sa != 0.0) {
#pragma unroll
        for (int r = 0; r < 3; ++r) Ff[r] = fa[r] + D(r, 0) * fa[0] + D(r, 1) * fa[1] + D(r, 2) * fa[2];
#pragma unroll
        for (int k = 0; k < 4; ++k) fg[k] = fa[0] * g[k][0] + fa[1] * g[k][1] + fa[2] * g[k][2];
    }
    // gradient (FP64): grad_a = V0 P g_a
    {
        M3<double> P = elem_stress(a, p, D, mu, lam, mk, mdir, mL);
        if (sa != 0.0)
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) P(r, c) += sa * Ff[r] * fa[c];
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const double gv = V0 * (P(r, 0) * g[k][0] + P(r, 1) * g[k][1] + P(r, 2) * g[k][2]);
                if (MODE == 2) {
                    if (a.ex_grad) a.ex_grad[(size_t)e * 12 + 3 * k + r] = gv;
                } else if (a.gint && dof_kept(a, 3 * nd[k] + r)) {
                    atomicAdd(a.gint + 3 * nd[k] + r, gv);
                }
            }
        if (MODE == 2 && a.ex_energy) {
            bool inv = false;
            a.ex_energy[e] = V0 * (elem_psi(a, D, mu, lam, &inv, true, mk, mdir, mL) +
                                   0.5 * sa * (Ff[0] * Ff[0] + Ff[1] * Ff[1] + Ff[2] * Ff[2] - 1.0));
        }
    }
    if (MODE == 1) return;
    // Hessian blocks in HT
    const HT w = (HT)(V0 * a.hscale);
    HT gh[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) gh[k][c] = (HT)g[k][c];
    const HT muh = (HT)mu, lamh = (HT)lam;
    if (gen) {
        // Stable Neo-Hookean and / or muscle: the 9x9 tangent, then its blocks
        M3<HT> Dh;
#pragma unroll
        for (int i = 0; i < 9; ++i) Dh.m[i] = (HT)D.m[i];
        Sym9<HT> S;
        elem_tangent<HT, PROJ>(a, Dh, muh, lamh, mk, mdir, mL, S);
        for (int la = 0; la < 4; ++la)
            for (int lb = 0; lb < 4; ++lb) {
                HT blk[9];
#pragma unroll
                for (int i = 0; i < 9; ++i) blk[i] = 0;
                block_from_tangent(S, gh[la], gh[lb], w, blk);
                if (sa != 0.0) {
                    const HT s_ab = w * (HT)(sa * fg[la] * fg[lb]);
                    blk[0] += s_ab;
                    blk[4] += s_ab;
                    blk[8] += s_ab;
                }
                if (MODE == 2) {
                    for (int c = 0; c < 3; ++c)
                        for (int d = 0; d < 3; ++d)
                            a.ex_hess[(size_t)e * 144 + (3 * la + c) * 12 + 3 * lb + d] = (double)blk[3 * c + d];
                } else {
                    scatter_block<HT>(a, e, la, lb, nd[la], blk);
                }
            }
    } else if (PROJ) {
        M3<HT> Dh;
#pragma unroll
        for (int i = 0; i < 9; ++i) Dh.m[i] = (HT)D.m[i];
        const NHPoint<HT> ph = nh_point(Dh);
        Sym9<HT> S;
        nh_projected_tangent(Dh, ph, muh, lamh, S);
#pragma unroll
        for (int la = 0; la < 4; ++la)
#pragma unroll
            for (int lb = 0; lb < 4; ++lb) {
                HT blk[9];
#pragma unroll
                for (int i = 0; i < 9; ++i) blk[i] = 0;
                block_from_tangent(S, gh[la], gh[lb], w, blk);
                if (sa != 0.0) {
                    const HT s_ab = w * (HT)(sa * fg[la] * fg[lb]);
                    blk[0] += s_ab;
                    blk[4] += s_ab;
                    blk[8] += s_ab;
                }
                if (MODE == 2) {
#pragma unroll
                    for (int c = 0; c < 3; ++c)
#pragma unroll
                        for (int d = 0; d < 3; ++d)
                            a.ex_hess[(size_t)e * 144 + (3 * la + c) * 12 + 3 * lb + d] = (double)blk[3 * c + d];
                } else {
                    scatter_block<HT>(a, e, la, lb, nd[la], blk);
                }
            }
    } else {
        // raw tangent: H_ab = V0 [mu (g_a.g_b) I + lam h_a h_b^T - c h_b h_a^T]
        const HT c = (HT)(lam * p.lnJ - mu);
        HT h[4][3];
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int r = 0; r < 3; ++r)
                h[k][r] = (HT)(p.FiT(r, 0) * g[k][0] + p.FiT(r, 1) * g[k][1] + p.FiT(r, 2) * g[k][2]);
#pragma unroll
        for (int la = 0; la < 4; ++la)
#pragma unroll
            for (int lb = 0; lb < 4; ++lb) {
                HT blk[9];
#pragma unroll
                for (int i = 0; i < 9; ++i) blk[i] = 0;
                block_raw(muh, lamh, c, gh[la], gh[lb], h[la], h[lb], w, blk);
                if (sa != 0.0) {
                    const HT s_ab = w * (HT)(sa * fg[la] * fg[lb]);
                    blk[0] += s_ab;
                    blk[4] += s_ab;
                    blk[8] += s_ab;
                }
                if (MODE == 2) {
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc)
#pragma unroll
                        for (int d = 0; d < 3; ++d)
                            a.ex_hess[(size_t)e * 144 + (3 * la + cc) * 12 + 3 * lb + d] = (double)blk[3 * cc + d];
                } else {
                    scatter_block<HT>(a, e, la, lb, nd[la], blk);
                }
            }
    }
}


// ---- tet4 / hex8, two-phase assembly on the sliced node-block layout ------
// Phase 1 evaluates every element once and stores:
//   eg[4 elem_row(e, a, ed) + r]  the FP64 gradient row of local node a (r < 3;
//                          one double of padding so a row is two 16-byte
//                          loads), tiled by 64 elements (pattern.h),
//   eb[blk elem_row(e, 8 a + b, 64) + i]  hex8: its 3x3 block (a, b), i < 9,
//                          padded to blk values (a 16-byte multiple);
//   eb[blk elem_row(e, t(a, b), 10) + i]  tet4: the ten blocks a <= b only,
//                          t(a, b) = 4 a - a (a - 1) / 2 + b - a (the element
//                          Hessian is symmetric: block (b, a) is the
//                          transpose of block (a, b)).
// tet4: k_tet_elem, one lane per element.  hex8: k_hex_grad (one lane per
// element, FP64 gradient rows) and k_hex_rows (one lane per element row a,
// its eight blocks summed over the eight Gauss points).
// Phase 2, MODE 0 (k_slot_sums): one lane per slot of the sliced layout sums
// the element blocks listed for that slot (pattern.h: sptr / sidx, ascending
// element order; a tet4 entry may name its stored block transposed) in
// registers and stores the block once; the lane of a node's own slot also
// sums the node's gradient rows.  MODE 1
// (k_grad_gather): one lane per node sums its gradient rows.  No atomics, no
// read-modify-write of the layout, every value written by exactly one lane,
// fixed-order sums.  MODE 0: gradient + Hessian; MODE 1: gradient only.
// A stored 3x3 element block: nine values and one pad, so every block starts
// 8-byte aligned (FP32, 40 B) or 16-byte aligned (FP64, 80 B)
template <typename HT>
constexpr int tet_blk() { return 10; }
// the unit a stored block is moved in: 8 bytes (FP32), 16 bytes (FP64)
template <typename HT>
using BlkPiece = typename std::conditional<sizeof(HT) == 4, uint2, uint4>::type;
// stored blocks per element: tet4 keeps the upper triangle
constexpr int elem_blocks(int ed) { return ed == 4 ? 10 : ed * ed; }

// GEN = 1: Stable Neo-Hookean and / or SoftCon muscle (the tangent is formed
// as a 9x9 and contracted); GEN = 0: plain Neo-Hookean (structured blocks)
template <typename HT, int PROJ, int MODE, int GEN>
__global__ __launch_bounds__(kBlock) void k_tet_elem(ElemArgs a, double* eg, HT* eb) {
    const int e = xcd_block() * blockDim.x + threadIdx.x;
    if (e >= a.E) return;
    double X[4][3], x[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int nd = a.conn[(size_t)k * a.E + e];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            X[k][c] = a.rest[3 * nd + c];
            x[k][c] = a.q[3 * nd + c];
        }
    }
    M3<double> Dm;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) Dm(r, c) = X[c][r] - X[3][r];
    const M3<double> Cm = cofactor3(Dm);
    const double detm = Dm.m[0] * Cm.m[0] + Dm.m[1] * Cm.m[1] + Dm.m[2] * Cm.m[2];
    const double V0 = fabs(detm) / 6.0;
    double g[4][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) g[k][c] = Cm(c, k) / detm;
#pragma unroll
    for (int c = 0; c < 3; ++c) g[3][c] = -(g[0][c] + g[1][c] + g[2][c]);
    M3<double> D;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double acc = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) acc += (x[k][r] - X[k][r]) * g[k][c];
            D(r, c) = acc;
        }
    const double mu = a.mu[e], lam = a.lam[e];
    const NHPoint<double> p = nh_point(D);
    if (!(p.J > 0) && !(GEN && snh_material(a))) atomicOr(a.flag, 1);
    double mdir[3] = {0.0, 0.0, 0.0}, mL = 0.0;
    const double mk = GEN ? muscle_of(a, e, mdir, &mL) : 0.0;
    double fa[3] = {0.0, 0.0, 0.0}, Ff[3] = {0.0, 0.0, 0.0}, fg[4] = {0.0, 0.0, 0.0, 0.0};
    const double sa = active_level(a, e, fa);
    if (sa != 0.0) {
#pragma unroll
        for (int r = 0; r < 3; ++r) Ff[r] = fa[r] + D(r, 0) * fa[0] + D(r, 1) * fa[1] + D(r, 2) * fa[2];
#pragma unroll
        for (int k = 0; k < 4; ++k) fg[k] = fa[0] * g[k][0] + fa[1] * g[k][1] + fa[2] * g[k][2];
    }
    {
        M3<double> P = GEN ? elem_stress(a, p, D, mu, lam, mk, mdir, mL) : nh_stress(p, mu, lam);
        if (sa != 0.0)
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) P(r, c) += sa * Ff[r] * fa[c];
#pragma unroll
        for (int la = 0; la < 4; ++la) {
            double gr[3];
#pragma unroll
            for (int r = 0; r < 3; ++r) gr[r] = V0 * (P(r, 0) * g[la][0] + P(r, 1) * g[la][1] + P(r, 2) * g[la][2]);
            // the padded 4-double row as two 16-byte stores
            double2* ge = reinterpret_cast<double2*>(eg + 4 * elem_row(e, la, 4));
            ge[0] = make_double2(gr[0], gr[1]);
            ge[1] = make_double2(gr[2], 0.0);
        }
    }
    if (MODE == 1) return;
    const HT w = (HT)(V0 * a.hscale);
    HT gh[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) gh[k][c] = (HT)g[k][c];
    const HT muh = (HT)mu, lamh = (HT)lam;
    Sym9<HT> S;
    HT cc = 0;
    HT h[4][3];
    if (GEN) {
        M3<HT> Dh;
#pragma unroll
        for (int i = 0; i < 9; ++i) Dh.m[i] = (HT)D.m[i];
        elem_tangent<HT, PROJ>(a, Dh, muh, lamh, mk, mdir, mL, S);
    } else if (PROJ) {
        M3<HT> Dh;
#pragma unroll
        for (int i = 0; i < 9; ++i) Dh.m[i] = (HT)D.m[i];
        const NHPoint<HT> ph = nh_point(Dh);
        nh_projected_tangent(Dh, ph, muh, lamh, S);
    } else {
        cc = (HT)(lam * p.lnJ - mu);
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int r = 0; r < 3; ++r)
                h[k][r] = (HT)(p.FiT(r, 0) * g[k][0] + p.FiT(r, 1) * g[k][1] + p.FiT(r, 2) * g[k][2]);
    }
    int t = 0;  // t(la, lb): row-major upper triangle
#pragma unroll
    for (int la = 0; la < 4; ++la)
#pragma unroll
        for (int lb = la; lb < 4; ++lb) {
            HT blk[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) blk[i] = 0;
            if (PROJ || GEN)
                block_from_tangent(S, gh[la], gh[lb], w, blk);
            else
                block_raw(muh, lamh, cc, gh[la], gh[lb], h[la], h[lb], w, blk);
            if (sa != 0.0) {
                const HT s_ab = w * (HT)(sa * fg[la] * fg[lb]);
                blk[0] += s_ab;
                blk[4] += s_ab;
                blk[8] += s_ab;
            }
            HT* dst = eb + (size_t)tet_blk<HT>() * elem_row(e, t, 10);
            if constexpr (sizeof(HT) == 4) {
                // the padded 10-float block as five 8-byte stores
                float2* d2 = reinterpret_cast<float2*>(dst);
                d2[0] = make_float2(blk[0], blk[1]);
                d2[1] = make_float2(blk[2], blk[3]);
                d2[2] = make_float2(blk[4], blk[5]);
                d2[3] = make_float2(blk[6], blk[7]);
                d2[4] = make_float2(blk[8], 0.0f);
            } else {
                // the padded 10-double block as five 16-byte stores
                double2* d2 = reinterpret_cast<double2*>(dst);
                d2[0] = make_double2(blk[0], blk[1]);
                d2[1] = make_double2(blk[2], blk[3]);
                d2[2] = make_double2(blk[4], blk[5]);
                d2[3] = make_double2(blk[6], blk[7]);
                d2[4] = make_double2(blk[8], 0.0);
            }
            ++t;
        }
}

// sum of node v's gradient rows over its incident (element, local node)
// codes ed e + a, in element order
__device__ __forceinline__ void node_gradient(const ElemArgs& a, int v, const int* incptr, const int* inc,
                                              const double* eg) {
    bool kept[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) kept[r] = dof_kept(a, 3 * v + r);
    double gacc[3] = {0.0, 0.0, 0.0};
    // incidence code ed e + a -> stored gradient row (ed is 4 or 8)
    const int lg = a.ed == 8 ? 3 : 2;
    auto row_of = [&](int code) { return elem_row(code >> lg, code & (a.ed - 1), a.ed); };
    const int j0 = incptr[v], j1 = incptr[v + 1];
    int j = j0;
    for (; j + 4 <= j1; j += 4) {  // four rows in flight, summed in list order
        int code[4];
        double2 g01[4], g2[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) code[u] = inc[j + u];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double2* ge = (const double2*)(eg + 4 * row_of(code[u]));
            g01[u] = ge[0];
            g2[u] = ge[1];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (kept[0]) gacc[0] += g01[u].x;
            if (kept[1]) gacc[1] += g01[u].y;
            if (kept[2]) gacc[2] += g2[u].x;
        }
    }
    for (; j < j1; ++j) {
        const double2* ge = (const double2*)(eg + 4 * row_of(inc[j]));
        const double2 g01 = ge[0], g2 = ge[1];
        if (kept[0]) gacc[0] += g01.x;
        if (kept[1]) gacc[1] += g01.y;
        if (kept[2]) gacc[2] += g2.x;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) a.gint[3 * v + r] = gacc[r];
}

__global__ __launch_bounds__(kBlock) void k_grad_gather(ElemArgs a, const int* incptr, const int* inc,
                                                        const double* eg) {
    const int v = xcd_block() * blockDim.x + threadIdx.x;
    if (v >= a.N) return;
    node_gradient(a, v, incptr, inc, eg);
}

// one lane per slot position p = sbase[s] + 64 k + lane (groups of 64 = one
// wavefront = slot k of slice s).  SYM (tet4): a list entry is 2 j + f, the
// stored block j, transposed when f = 1; otherwise the entry is the block.
template <typename HT, bool SYM>
__global__ __launch_bounds__(kBlock) void k_slot_sums(ElemArgs a, int nq, const int* qslice, const int* sptr,
                                                      const int* sidx, const HT* eb, const int* selfslot,
                                                      const int* incptr, const int* inc, const double* eg) {
    const int p = xcd_block() * blockDim.x + threadIdx.x;
    const int q = p / kSlice, lane = p % kSlice;
    if (q >= nq) return;
    const int s = qslice[q];
    const int k = q - a.sbase[s] / kSlice;
    constexpr int kPieces = tet_blk<HT>() * (int)sizeof(HT) / (int)sizeof(BlkPiece<HT>);
    HT acc[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) acc[i] = (HT)0;
    union Blk {
        BlkPiece<HT> q4[kPieces];
        HT v[tet_blk<HT>()];
    };
    auto add = [&](const Blk& b, int code) {
        const bool tr = SYM && (code & 1);
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[i] += tr ? b.v[3 * (i % 3) + i / 3] : b.v[i];
    };
    auto block_of = [](int code) { return SYM ? code >> 1 : code; };
    const int c0 = sptr[p], c1 = sptr[p + 1];
    // groups of up to kGroup list entries (a typical list is one group): their
    // indices, then their blocks, are loaded together; the sums keep the list order
    constexpr int kGroup = sizeof(HT) == 4 ? 8 : 4;
    for (int c = c0; c < c1; c += kGroup) {
        const int m = min(kGroup, c1 - c);
        int id[kGroup];
        Blk blk[kGroup];
#pragma unroll
        for (int u = 0; u < kGroup; ++u) id[u] = u < m ? sidx[c + u] : 0;
#pragma unroll
        for (int u = 0; u < kGroup; ++u)
            if (u < m) {
                const BlkPiece<HT>* src = (const BlkPiece<HT>*)(eb + (size_t)tet_blk<HT>() * block_of(id[u]));
#pragma unroll
                for (int i = 0; i < kPieces; ++i) blk[u].q4[i] = src[i];
            }
#pragma unroll
        for (int u = 0; u < kGroup; ++u)
            if (u < m) add(blk[u], id[u]);
    }
    HT* vals = (HT*)a.vals;
    const long long vb = 9LL * a.sbase[s] + (long long)kSlice * 9 * k + lane;
#pragma unroll
    for (int i = 0; i < 9; ++i) vals[vb + (long long)kSlice * i] = acc[i];
    const int v = s * kSlice + lane;
    if (v < a.N && k == selfslot[v]) node_gradient(a, v, incptr, inc, eg);
}

// ---- hex8: one element per lane --------------------------------------------
// Reference cube corners (hexahedron.cpp:171-179)
__constant__ signed char kHexXi[8][3] = {{1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, -1},
                                         {1, -1, 1},  {1, 1, 1},  {-1, 1, 1},  {-1, -1, 1}};

// G_s = hatG_s hatF_s^-1 with hatF_s = X^T hatG_s (hexahedron.cpp:190-205)
template <typename T>
__device__ __forceinline__ void hex_shape_grads(const double X[8][3], int s, T G[8][3]) {
    const double gq = 0.57735026918962576451;  // 1/sqrt(3)
    const double qs0 = gq * kHexXi[s][0], qs1 = gq * kHexXi[s][1], qs2 = gq * kHexXi[s][2];
    // reference-cube gradient of shape function i at Gauss point s (recomputed
    // where used: a few flops, and no 8x3 FP64 array kept live)
    auto hg = [&](int i, double h[3]) {
        const double n0 = kHexXi[i][0], n1 = kHexXi[i][1], n2 = kHexXi[i][2];
        h[0] = 0.125 * (n0 * (1 + n1 * qs1) * (1 + n2 * qs2));
        h[1] = 0.125 * (n1 * (1 + n0 * qs0) * (1 + n2 * qs2));
        h[2] = 0.125 * (n2 * (1 + n0 * qs0) * (1 + n1 * qs1));
    };
    M3<double> hF;
#pragma unroll
    for (int k = 0; k < 9; ++k) hF.m[k] = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        double h[3];
        hg(i, h);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) hF(r, c) += X[i][r] * h[c];
    }
    const M3<double> C = cofactor3(hF);
    const double det = hF.m[0] * C.m[0] + hF.m[1] * C.m[1] + hF.m[2] * C.m[2];
    // hF^-1 (k, c) = C(c, k) / det
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        double h[3];
        hg(i, h);
#pragma unroll
        for (int c = 0; c < 3; ++c) G[i][c] = (T)((h[0] * C(c, 0) + h[1] * C(c, 1) + h[2] * C(c, 2)) / det);
    }
}

__device__ __forceinline__ double tet_vol(const double* a, const double* b, const double* c, const double* d) {
    const double ab0 = b[0] - a[0], ab1 = b[1] - a[1], ab2 = b[2] - a[2];
    const double ac0 = c[0] - a[0], ac1 = c[1] - a[1], ac2 = c[2] - a[2];
    const double ad0 = d[0] - a[0], ad1 = d[1] - a[1], ad2 = d[2] - a[2];
    return fabs(ab0 * (ac1 * ad2 - ac2 * ad1) + ab1 * (ac2 * ad0 - ac0 * ad2) + ab2 * (ac0 * ad1 - ac1 * ad0)) / 6.0;
}

// V0 of a hex: the five-tet split {0,1,3,4} {1,2,3,6} {1,4,5,6} {3,4,6,7}
// {1,3,4,6} (hexahedron.cpp:106-130), summed in that order.  Literal
// indices keep X in registers (a table lookup would index it at run time).
__device__ __forceinline__ double hex_volume(const double X[8][3]) {
    double V0 = tet_vol(X[0], X[1], X[3], X[4]);
    V0 += tet_vol(X[1], X[2], X[3], X[6]);
    V0 += tet_vol(X[1], X[4], X[5], X[6]);
    V0 += tet_vol(X[3], X[4], X[6], X[7]);
    V0 += tet_vol(X[1], X[3], X[4], X[6]);
    return V0;
}

template <typename HT, int PROJ, int MODE>
__global__ __launch_bounds__(128) void k_hex(ElemArgs a) {
    const int e = xcd_block() * blockDim.x + threadIdx.x;
    if (e >= a.E) return;
    int nd[8];
    double X[8][3], u[8][3];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        nd[k] = a.conn[(size_t)k * a.E + e];
        double xk[3];
        gather_node(a, nd[k], X[k], xk);
#pragma unroll
        for (int c = 0; c < 3; ++c) u[k][c] = xk[c] - X[k][c];
    }
    const double V0 = hex_volume(X);
    const double mu = a.mu[e], lam = a.lam[e];
    double mdir[3] = {0.0, 0.0, 0.0}, mL = 0.0;
    const double mk = muscle_of(a, e, mdir, &mL);
    const bool gen = snh_material(a) || mk != 0.0;
    const double wq = V0 * 0.125;  // quadrature weight 1/8 times V0 (neoHookean.cpp:26-27)
    // gradient and energy, FP64
    {
        double gr[8][3];
#pragma unroll
        for (int k = 0; k < 8; ++k) gr[k][0] = gr[k][1] = gr[k][2] = 0;
        double en = 0;
        for (int s = 0; s < 8; ++s) {
            double G[8][3];
            hex_shape_grads<double>(X, s, G);
            M3<double> D;
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    double acc = 0;
#pragma unroll
                    for (int k = 0; k < 8; ++k) acc += u[k][r] * G[k][c];
                    D(r, c) = acc;
                }
            const NHPoint<double> p = nh_point(D);
            if (!(p.J > 0) && !snh_material(a)) atomicOr(a.flag, 1);
            const M3<double> P = elem_stress(a, p, D, mu, lam, mk, mdir, mL);
#pragma unroll
            for (int k = 0; k < 8; ++k)
#pragma unroll
                for (int r = 0; r < 3; ++r) gr[k][r] += wq * (P(r, 0) * G[k][0] + P(r, 1) * G[k][1] + P(r, 2) * G[k][2]);
            if (MODE == 2) {
                bool inv = false;
                en += wq * elem_psi(a, D, mu, lam, &inv, true, mk, mdir, mL);
            }
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                if (MODE == 2) {
                    if (a.ex_grad) a.ex_grad[(size_t)e * 24 + 3 * k + r] = gr[k][r];
                } else if (a.gint && dof_kept(a, 3 * nd[k] + r)) {
                    atomicAdd(a.gint + 3 * nd[k] + r, gr[k][r]);
                }
            }
        if (MODE == 2 && a.ex_energy) a.ex_energy[e] = en;
    }
    if (MODE == 1) return;
    // Hessian: row block la at a time, Gauss points recomputed per row block
    const HT muh = (HT)mu, lamh = (HT)lam;
    const HT w = (HT)(wq * a.hscale);
    for (int la = 0; la < 8; ++la) {
        HT acc[8][9];
#pragma unroll
        for (int lb = 0; lb < 8; ++lb)
#pragma unroll
            for (int i = 0; i < 9; ++i) acc[lb][i] = 0;
        for (int s = 0; s < 8; ++s) {
            HT G[8][3];
            hex_shape_grads<HT>(X, s, G);
            M3<HT> D;
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    HT t = 0;
#pragma unroll
                    for (int k = 0; k < 8; ++k) t += (HT)u[k][r] * G[k][c];
                    D(r, c) = t;
                }
            const NHPoint<HT> p = nh_point(D);
            if (gen) {
                Sym9<HT> S;
                elem_tangent<HT, PROJ>(a, D, muh, lamh, mk, mdir, mL, S);
                for (int lb = 0; lb < 8; ++lb) block_from_tangent(S, G[la], G[lb], w, acc[lb]);
            } else if (PROJ) {
                Sym9<HT> S;
                nh_projected_tangent(D, p, muh, lamh, S);
#pragma unroll
                for (int lb = 0; lb < 8; ++lb) block_from_tangent(S, G[la], G[lb], w, acc[lb]);
            } else {
                const HT c = lamh * p.lnJ - muh;
                HT ha[3];
#pragma unroll
                for (int r = 0; r < 3; ++r) ha[r] = p.FiT(r, 0) * G[la][0] + p.FiT(r, 1) * G[la][1] + p.FiT(r, 2) * G[la][2];
#pragma unroll
                for (int lb = 0; lb < 8; ++lb) {
                    HT hb[3];
#pragma unroll
                    for (int r = 0; r < 3; ++r)
                        hb[r] = p.FiT(r, 0) * G[lb][0] + p.FiT(r, 1) * G[lb][1] + p.FiT(r, 2) * G[lb][2];
                    block_raw(muh, lamh, c, G[la], G[lb], ha, hb, w, acc[lb]);
                }
            }
        }
#pragma unroll
        for (int lb = 0; lb < 8; ++lb) {
            if (MODE == 2) {
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int d = 0; d < 3; ++d)
                        a.ex_hess[(size_t)e * 576 + (3 * la + c) * 24 + 3 * lb + d] = (double)acc[lb][3 * c + d];
            } else {
                scatter_block<HT>(a, e, la, lb, nd[la], acc[lb]);
            }
        }
    }
}

// ---- hex8 phase 1 of the two-phase assembly ------------------------------
template <typename UT>
__device__ __forceinline__ double hex_load(const ElemArgs& a, int e, double X[8][3], UT u[8][3]) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int nd = a.conn[(size_t)k * a.E + e];
        double xk[3];
        gather_node(a, nd, X[k], xk);
#pragma unroll
        for (int c = 0; c < 3; ++c) u[k][c] = (UT)(xk[c] - X[k][c]);
    }
    const double V0 = hex_volume(X);
    return V0 * 0.125;  // quadrature weight 1/8 times V0 (neoHookean.cpp:26-27)
}

// FP64 gradient rows of one element (same sums as k_hex); GEN as in k_tet_elem
template <int GEN>
__global__ __launch_bounds__(128) void k_hex_grad(ElemArgs a, double* eg) {
    const int e = xcd_block() * blockDim.x + threadIdx.x;
    if (e >= a.E) return;
    double X[8][3], u[8][3];
    const double wq = hex_load(a, e, X, u);
    const double mu = a.mu[e], lam = a.lam[e];
    double mdir[3] = {0.0, 0.0, 0.0}, mL = 0.0;
    const double mk = GEN ? muscle_of(a, e, mdir, &mL) : 0.0;
    double gr[8][3];
#pragma unroll
    for (int k = 0; k < 8; ++k) gr[k][0] = gr[k][1] = gr[k][2] = 0;
    for (int s = 0; s < 8; ++s) {
        double G[8][3];
        hex_shape_grads<double>(X, s, G);
        M3<double> D;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double acc = 0;
#pragma unroll
                for (int k = 0; k < 8; ++k) acc += u[k][r] * G[k][c];
                D(r, c) = acc;
            }
        const NHPoint<double> p = nh_point(D);
        if (!(p.J > 0) && !(GEN && snh_material(a))) atomicOr(a.flag, 1);
        const M3<double> P = GEN ? elem_stress(a, p, D, mu, lam, mk, mdir, mL) : nh_stress(p, mu, lam);
#pragma unroll
        for (int k = 0; k < 8; ++k)
#pragma unroll
            for (int r = 0; r < 3; ++r) gr[k][r] += wq * (P(r, 0) * G[k][0] + P(r, 1) * G[k][1] + P(r, 2) * G[k][2]);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
        for (int r = 0; r < 3; ++r) eg[4 * elem_row(e, k, 8) + r] = gr[k][r];
}

// Hessian rows of hex8 elements: eight lanes per element.  First lane j
// evaluates Gauss point j (shape gradients, F, the projected 9x9 tangent, or
// F^-T and lambda ln J - mu for the raw tangent) into LDS; then lane j builds
// row j, its eight blocks summed over the Gauss points in order.  Each
// Gauss-point tangent is evaluated once per element, and t_i = g_a^T S(3i+.,
// .) is formed once per Gauss point and reused for the eight blocks.
constexpr int kHexRowThreads = 128;  // 16 elements per workgroup
template <typename HT, int PROJ, int GEN>
__global__ __launch_bounds__(kHexRowThreads) void k_hex_rows(ElemArgs a, HT* eb) {
    constexpr int kEl = kHexRowThreads / 8;
    __shared__ HT sS[kEl][8][45];  // per Gauss point: Sym9 (PROJ) or F^-T and c (raw)
    __shared__ HT sG[kEl][8][24];  // per Gauss point: shape gradients G[k][c]
    const int le = threadIdx.x >> 3, j = threadIdx.x & 7;
    const int e = (xcd_block() * kHexRowThreads + threadIdx.x) >> 3;
    const bool live = e < a.E;
    HT w = 0;
    if (live) {
        double X[8][3];
        HT u[8][3];
        const double wq = hex_load(a, e, X, u);
        const double mu = a.mu[e], lam = a.lam[e];
        w = (HT)(wq * a.hscale);
        HT G[8][3];
        hex_shape_grads<HT>(X, j, G);
        M3<HT> D;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                HT tt = 0;
#pragma unroll
                for (int k = 0; k < 8; ++k) tt += u[k][r] * G[k][c];
                D(r, c) = tt;
            }
        const NHPoint<HT> p = nh_point(D);
        if (GEN) {
            double mdir[3] = {0.0, 0.0, 0.0}, mL = 0.0;
            const double mk = muscle_of(a, e, mdir, &mL);
            Sym9<HT> S;
            elem_tangent<HT, PROJ>(a, D, (HT)mu, (HT)lam, mk, mdir, mL, S);
#pragma unroll
            for (int i = 0; i < 45; ++i) sS[le][j][i] = S.a[i];
        } else if (PROJ) {
            Sym9<HT> S;
            nh_projected_tangent(D, p, (HT)mu, (HT)lam, S);
#pragma unroll
            for (int i = 0; i < 45; ++i) sS[le][j][i] = S.a[i];
        } else {
#pragma unroll
            for (int i = 0; i < 9; ++i) sS[le][j][i] = p.FiT.m[i];
            sS[le][j][9] = (HT)lam * p.lnJ - (HT)mu;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c) sG[le][j][3 * k + c] = G[k][c];
    }
    __syncthreads();
    if (!live) return;
    const int la = j;
    HT acc[8][9];
#pragma unroll
    for (int lb = 0; lb < 8; ++lb)
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[lb][i] = 0;
    for (int s = 0; s < 8; ++s) {
        const HT* Gs = sG[le][s];
        const HT ga[3] = {Gs[3 * la], Gs[3 * la + 1], Gs[3 * la + 2]};
        if (PROJ || GEN) {
            Sym9<HT> S;
#pragma unroll
            for (int i = 0; i < 45; ++i) S.a[i] = sS[le][s][i];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                HT t[9];  // t[(k,l)] = sum_j ga_j S(3i+j, 3k+l)
#pragma unroll
                for (int kl = 0; kl < 9; ++kl) {
                    HT tacc = 0;
#pragma unroll
                    for (int jj = 0; jj < 3; ++jj) tacc += ga[jj] * S.get(3 * i + jj, kl);
                    t[kl] = tacc;
                }
#pragma unroll
                for (int lb = 0; lb < 8; ++lb)
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        HT bacc = 0;
#pragma unroll
                        for (int l = 0; l < 3; ++l) bacc += t[3 * k + l] * Gs[3 * lb + l];
                        acc[lb][3 * i + k] += w * bacc;
                    }
            }
        } else {
            M3<HT> FiT;
#pragma unroll
            for (int i = 0; i < 9; ++i) FiT.m[i] = sS[le][s][i];
            const HT c = sS[le][s][9];
            HT ha[3];
#pragma unroll
            for (int r = 0; r < 3; ++r) ha[r] = FiT(r, 0) * ga[0] + FiT(r, 1) * ga[1] + FiT(r, 2) * ga[2];
            const HT muh = (HT)a.mu[e], lamh = (HT)a.lam[e];
#pragma unroll
            for (int lb = 0; lb < 8; ++lb) {
                const HT gb[3] = {Gs[3 * lb], Gs[3 * lb + 1], Gs[3 * lb + 2]};
                HT hb[3];
#pragma unroll
                for (int r = 0; r < 3; ++r) hb[r] = FiT(r, 0) * gb[0] + FiT(r, 1) * gb[1] + FiT(r, 2) * gb[2];
                block_raw(muh, lamh, c, ga, gb, ha, hb, w, acc[lb]);
            }
        }
    }
#pragma unroll
    for (int lb = 0; lb < 8; ++lb)
#pragma unroll
        for (int i = 0; i < 9; ++i) eb[tet_blk<HT>() * elem_row(e, 8 * la + lb, 64) + i] = acc[lb][i];
}

// ---- element energy: absolute (diff = 0) or E(q + alpha dq) - E(q) -------
// absolute: include Stable Neo-Hookean's rest energy (left out of differences)
__device__ double tet_energy_at(const ElemArgs& a, int e, const int nd[4], double alpha, bool* inv, bool absolute) {
    double X[4][3], x[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            X[k][c] = a.rest[3 * nd[k] + c];
            double v = a.q[3 * nd[k] + c];
            if (alpha != 0.0) v += alpha * a.dq[3 * nd[k] + c];
            x[k][c] = v;
        }
    if (a.material == SORS_TEST_POLY) {
        double en = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c) en += poly_psi(x[k][c]);
        return en;
    }
    M3<double> Dm;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) Dm(r, c) = X[c][r] - X[3][r];
    const M3<double> Cm = cofactor3(Dm);
    const double detm = Dm.m[0] * Cm.m[0] + Dm.m[1] * Cm.m[1] + Dm.m[2] * Cm.m[2];
    double g[4][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) g[k][c] = Cm(c, k) / detm;
#pragma unroll
    for (int c = 0; c < 3; ++c) g[3][c] = -(g[0][c] + g[1][c] + g[2][c]);
    M3<double> D;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double acc = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) acc += (x[k][r] - X[k][r]) * g[k][c];
            D(r, c) = acc;
        }
    double mdir[3] = {0.0, 0.0, 0.0}, mL = 0.0;
    const double mk = muscle_of(a, e, mdir, &mL);
    double en = elem_psi(a, D, a.mu[e], a.lam[e], inv, absolute, mk, mdir, mL);
    double fa[3];
    const double sa = active_level(a, e, fa);
    if (sa != 0.0) {
        double ff = 0.0;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const double v = fa[r] + D(r, 0) * fa[0] + D(r, 1) * fa[1] + D(r, 2) * fa[2];
            ff += v * v;
        }
        en += 0.5 * sa * (ff - 1.0);
    }
    return fabs(detm) / 6.0 * en;
}

__device__ double hex_energy_at(const ElemArgs& a, int e, const int nd[8], double alpha, bool* inv, bool absolute) {
    double X[8][3], u[8][3];
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            X[k][c] = a.rest[3 * nd[k] + c];
            double v = a.q[3 * nd[k] + c];
            if (alpha != 0.0) v += alpha * a.dq[3 * nd[k] + c];
            u[k][c] = v - X[k][c];
        }
    const double V0 = hex_volume(X);
    double mdir[3] = {0.0, 0.0, 0.0}, mL = 0.0;
    const double mk = muscle_of(a, e, mdir, &mL);
    double en = 0;
    for (int s = 0; s < 8; ++s) {
        double G[8][3];
        hex_shape_grads<double>(X, s, G);
        M3<double> D;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double acc = 0;
#pragma unroll
                for (int k = 0; k < 8; ++k) acc += u[k][r] * G[k][c];
                D(r, c) = acc;
            }
        en += elem_psi(a, D, a.mu[e], a.lam[e], inv, absolute, mk, mdir, mL);
    }
    return V0 * 0.125 * en;
}

template <int ED>
__global__ __launch_bounds__(kBlock) void k_elem_energy(ElemArgs a, int diff, double* out, Reduce red) {
    double v[1] = {0.0};
    bool inv = false;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < a.E; e += gridDim.x * blockDim.x) {
        int nd[ED];
#pragma unroll
        for (int k = 0; k < ED; ++k) nd[k] = a.conn[(size_t)k * a.E + e];
        double en;
        if (ED == 4) {
            en = tet_energy_at(a, e, nd, a.alpha, &inv, !diff);
            if (diff) en -= tet_energy_at(a, e, nd, 0.0, &inv, false);
        } else {
            en = hex_energy_at(a, e, nd, a.alpha, &inv, !diff);
            if (diff) en -= hex_energy_at(a, e, nd, 0.0, &inv, false);
        }
        v[0] += en;
    }
    if (inv) atomicOr(a.flag, 2);
    block_partial<1>(v, red.partials);
}

// ===========================================================================
// CFL time step (Solver::calculate_CFL_timestep, solver.cpp:300-331): the
// shortest distance between two vertices of one element at positions q
// (Element::compute_minimal_edge_length, element.h:131-142: every vertex
// pair, so a hex's face and body diagonals too) and the largest |v_prev|
// component.  One (min, max) pair per block in part[b], part[kMaxRedBlocks + b].
// ===========================================================================
template <bool MIN>
__device__ __forceinline__ double block_extremum(double v) {
    __shared__ double sh[kBlock / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = MIN ? fmin(v, o) : fmax(v, o);
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) v = MIN ? fmin(v, sh[w]) : fmax(v, sh[w]);
    __syncthreads();
    return v;  // valid in thread 0
}

template <int ED>
__global__ __launch_bounds__(kBlock) void k_cfl_partials(int E, const int* conn, const double* q, int n,
                                                         const double* v, double* part) {
    double emin = INFINITY, vmax = 0.0;
    const int stride = gridDim.x * blockDim.x;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += stride) {
        double x[ED][3];
#pragma unroll
        for (int k = 0; k < ED; ++k) {
            const int nd = conn[(size_t)k * E + e];
#pragma unroll
            for (int c = 0; c < 3; ++c) x[k][c] = q[3 * nd + c];
        }
#pragma unroll
        for (int i = 0; i < ED; ++i)
#pragma unroll
            for (int j = i + 1; j < ED; ++j) {
                const double d0 = x[i][0] - x[j][0], d1 = x[i][1] - x[j][1], d2 = x[i][2] - x[j][2];
                emin = fmin(emin, sqrt(d0 * d0 + d1 * d1 + d2 * d2));
            }
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) vmax = fmax(vmax, fabs(v[i]));
    emin = block_extremum<true>(emin);
    vmax = block_extremum<false>(vmax);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = emin;
        part[kMaxRedBlocks + blockIdx.x] = vmax;
    }
}

// ===========================================================================
// PressureForce (pressureForce.cpp:14-185): per face the area vector A (a
// triangle's (x1 - x0) x (x2 - x0) / 2; a quad's triangles (x0, x1, x3) and
// (x2, x3, x1)) and the force -p A / nv on each of its vertices; per node the
// sum over its faces in ascending order (no atomics).  The Jacobian blocks
// (row vertex j, column vertex k) = p/6 hessX[k], added row by row into the
// node-block layout (one lane per row node).
// ===========================================================================
__device__ __forceinline__ void press_pos(const double* q, const double* dq, double alpha, int v, double x[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double t = q[3 * v + k];
        if (dq) t += alpha * dq[3 * v + k];
        x[k] = t;
    }
}

__device__ __forceinline__ void half_cross(const double a[3], const double b[3], double out[3]) {
    out[0] = 0.5 * (a[1] * b[2] - a[2] * b[1]);
    out[1] = 0.5 * (a[2] * b[0] - a[0] * b[2]);
    out[2] = 0.5 * (a[0] * b[1] - a[1] * b[0]);
}

// ff[3 f + k] = -p A_k / nv at x = q + alpha dq (dq may be null)
__global__ __launch_bounds__(kBlock) void k_pressure_faces(PressArgs a, const double* q, const double* dq,
                                                           double alpha, double* ff) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= a.nf) return;
    double x[4][3];
    for (int j = 0; j < a.nv; ++j) press_pos(q, dq, alpha, a.faces[(size_t)j * a.nf + f], x[j]);
    double A[3], u[3], v[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        u[k] = x[1][k] - x[0][k];
        v[k] = x[a.nv == 3 ? 2 : 3][k] - x[0][k];
    }
    half_cross(u, v, A);
    if (a.nv == 4) {
        double B[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            u[k] = x[3][k] - x[2][k];
            v[k] = x[1][k] - x[2][k];
        }
        half_cross(u, v, B);
#pragma unroll
        for (int k = 0; k < 3; ++k) A[k] += B[k];
    }
    const double p = a.pact[a.group[f]];
#pragma unroll
    for (int k = 0; k < 3; ++k) ff[3 * (size_t)f + k] = -p * A[k] / a.nv;
}

// fp[3 v + k] = sum over the faces of v, ascending
__global__ __launch_bounds__(kBlock) void k_pressure_nodes(PressArgs a, const double* ff, double* fp) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.N) return;
    double s[3] = {0.0, 0.0, 0.0};
    for (int t = a.incptr[v]; t < a.incptr[v + 1]; ++t) {
        const int f = a.inc[t];
#pragma unroll
        for (int k = 0; k < 3; ++k) s[k] += ff[3 * (size_t)f + k];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) fp[3 * (size_t)v + k] = s[k];
}

// the reference's 3x3 block for a vertex with the other two triangle
// vertices xa, xb: rows (0, xa2 - xb2, xb1 - xa1), (xb2 - xa2, 0, xa0 - xb0),
// (xa1 - xb1, xb0 - xa0, 0)
__device__ __forceinline__ void press_block(const double xa[3], const double xb[3], double M[9], bool add) {
    const double m[9] = {0.0, -xb[2] + xa[2], xb[1] - xa[1], xb[2] - xa[2], 0.0, -xb[0] + xa[0],
                         -xb[1] + xa[1], xb[0] - xa[0], 0.0};
#pragma unroll
    for (int t = 0; t < 9; ++t) M[t] = add ? M[t] + m[t] : m[t];
}

// rows of node v: += scale * p/6 hessX[k] for every face of v and vertex k of
// that face (both nodes free); *flag |= 4 when a face pair is not in the layout
template <typename HT>
__global__ __launch_bounds__(kBlock) void k_pressure_jacobian(PressArgs a, const double* q, const int* sbase,
                                                              const int* bcol, const int* self_slot, const int* mask,
                                                              double scale, HT* vals, int* flag) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.N) return;
    if (mask && mask[3 * v]) return;
    const int sl = v >> 6, lane = v & (kSlice - 1);
    const int width = (sbase[sl + 1] - sbase[sl]) / kSlice;
    for (int t = a.incptr[v]; t < a.incptr[v + 1]; ++t) {
        const int f = a.inc[t];
        int fv[4];
        double x[4][3];
        for (int j = 0; j < a.nv; ++j) {
            fv[j] = a.faces[(size_t)j * a.nf + f];
            press_pos(q, nullptr, 0.0, fv[j], x[j]);
        }
        double H[4][9];
        if (a.nv == 3) {
            press_block(x[1], x[2], H[0], false);
            press_block(x[2], x[0], H[1], false);
            press_block(x[0], x[1], H[2], false);
        } else {
            press_block(x[1], x[3], H[0], false);  // triangle (x0, x1, x3)
            press_block(x[3], x[0], H[1], false);
            press_block(x[0], x[1], H[3], false);
            press_block(x[3], x[1], H[2], false);  // triangle (x2, x3, x1)
            press_block(x[1], x[2], H[3], true);
            press_block(x[2], x[3], H[1], true);
        }
        const double c6 = a.pact[a.group[f]] / 6;
        for (int k = 0; k < a.nv; ++k) {
            const int cn = fv[k];
            if (mask && mask[3 * cn]) continue;
            int slot = -1;
            if (cn == v) {
                slot = self_slot[v];
            } else {
                for (int w = 0; w < width; ++w)
                    if (bcol[sbase[sl] + kSlice * w + lane] == cn) {
                        slot = w;
                        break;
                    }
            }
            if (slot < 0) {
                atomicOr(flag, 4);
                continue;
            }
#pragma unroll
            for (int l = 0; l < 3; ++l)
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    const long long idx = block_index(sbase, v, slot, l, m);
                    vals[idx] = (HT)((double)vals[idx] + scale * (c6 * H[k][3 * l + m]));
                }
        }
    }
}

// ===========================================================================
// Node kernels: inertia, gravity, damping, external force, Dirichlet rows
// (integrator.cpp:45-124, energy.cpp:282-306, 381-388)
// ===========================================================================
template <typename HT>
__global__ __launch_bounds__(kBlock) void k_node_terms(NodeArgs a, double* gg_out, Reduce red) {
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x) {
        const double m = a.mass[i], qi = a.q[i];
        double gi = a.gint ? a.gint[i] : 0.0;
        if (a.fext) gi -= a.fext[i];
        if (a.fpres) gi -= a.fpres[i];
        const double gc = -m * a.grav[i % 3];
        double g, hd = 0.0;
        if (a.dt != 0.0) {
            if (a.scheme == SORS_BACKWARD_EULER) {
                const double vel = (qi - a.qprev[i]) / a.dt;
                g = m * ((vel - a.vprev[i]) / a.dt) + gc + a.damping * m * vel + gi;
                hd = m / (a.dt * a.dt) + a.damping * m / a.dt;
            } else {
                const double vel = 2 * (qi - a.qprev[i]) / a.dt - a.vprev[i];
                g = m * ((vel - a.vprev[i]) / a.dt) + gc + a.damping * m * vel + 0.5 * (gi + a.gprev[i]);
                hd = 2 * m / (a.dt * a.dt) + 2 * a.damping * m / a.dt;
            }
            if (a.gtmp) a.gtmp[i] = gi;
        } else {
            g = gc + gi;
        }
        if (a.gcon) g += a.gcon[i];
        const bool dir = !a.sqp && a.mask && a.mask[i];
        if (dir) g = -(a.qprev[i] + a.dt * a.bcval[i] - qi);
        a.g[i] = g;
        v[0] += g * g;
        if (a.vals) {
            HT* vals = (HT*)a.vals;
            const int di = a.diag[i];
            if (di < 0) {
                // row without a diagonal slot: only empty rows of isolated nodes at dt == 0
            } else if (dir)
                vals[di] = (HT)1;
            else if (hd != 0.0)
                vals[di] += (HT)hd;
        }
    }
    block_partial<1>(v, red.partials);
}

// energy terms of the nodes (integrate_energy, integrator.cpp:3-43, and the
// external-force work, energy.cpp:228-244).  diff != 0: E(q + alpha dq) - E(q)
// in difference form (no cancellation of the absolute terms).
__global__ __launch_bounds__(kBlock) void k_node_energy(NodeArgs a, const double* dq, double alpha, int diff,
                                                         double* out, Reduce red) {
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x) {
        const double m = a.mass[i], gc = a.grav[i % 3];
        const double qi = a.q[i];
        double en;
        if (diff) {
            const double d = alpha * dq[i];
            en = -m * d * gc;
            if (a.fext) en += a.fext[i] * d;
            if (a.fpres) en += a.fpres1[i] * (qi + d - a.qprev[i]) - a.fpres[i] * (qi - a.qprev[i]);
            if (a.dt != 0.0) {
                if (a.scheme == SORS_BACKWARD_EULER) {
                    const double vel = (qi - a.qprev[i]) / a.dt, acc = (vel - a.vprev[i]) / a.dt;
                    const double dv = d / a.dt, da = dv / a.dt;
                    en += 0.5 * m * (2 * acc * da + da * da) * a.dt * a.dt;
                    en += 0.5 * a.damping * a.dt * m * (2 * vel * dv + dv * dv);
                } else {
                    const double vel = 2 * (qi - a.qprev[i]) / a.dt - a.vprev[i], acc = (vel - a.vprev[i]) / a.dt;
                    const double dv = 2 * d / a.dt, da = dv / a.dt;
                    en += 0.25 * m * (2 * acc * da + da * da) * a.dt * a.dt;
                    en += 0.25 * a.damping * a.dt * m * (2 * vel * dv + dv * dv);
                }
            }
        } else {
            en = -m * qi * gc;
            if (a.fext) en += a.fext[i] * (qi - a.qprev[i]);
            if (a.fpres) en += a.fpres[i] * (qi - a.qprev[i]);
            if (a.dt != 0.0) {
                if (a.scheme == SORS_BACKWARD_EULER) {
                    const double vel = (qi - a.qprev[i]) / a.dt, acc = (vel - a.vprev[i]) / a.dt;
                    en += 0.5 * m * acc * acc * a.dt * a.dt + 0.5 * a.damping * a.dt * m * vel * vel;
                } else {
                    const double vel = 2 * (qi - a.qprev[i]) / a.dt - a.vprev[i], acc = (vel - a.vprev[i]) / a.dt;
                    en += 0.5 * m * acc * acc * a.dt * a.dt / 2 + 0.25 * a.damping * a.dt * m * vel * vel;
                }
            }
        }
        v[0] += en;
    }
    block_partial<1>(v, red.partials);
}

// ===========================================================================
// Contact: penalty 1/2 k w d^2 against planes and disks (d < 0 penetrating;
// w = 1 for a plane, a C1 cubic rim falloff over kDiskRim outside a disk's
// radius) plus lagged Coulomb friction mu N f0(|uT|) with the C2 smoothing
// f0 of Li et al., "Incremental Potential Contact" (SIGGRAPH 2020, Sec. 5),
// eps = eps_v dt; N, the normal and the primitive's displacement are fixed
// at the start of the step.  Exact gradients and Hessian blocks; each
// node's 3x3 contact block is projected to PSD (it differs from the exact
// block only on a disk rim).  One lane per candidate node; the node's 3x3
// self block takes the Hessian.
// ===========================================================================
__device__ __forceinline__ double rim_weight_dev(double rho, double r, double* ds, double* d2s) {
    *ds = 0.0;
    *d2s = 0.0;
    if (rho <= r) return 1.0;
    const double z = (rho - r) / kDiskRim;
    if (z >= 1.0) return 0.0;
    *ds = (-6.0 * z + 6.0 * z * z) / kDiskRim;
    *d2s = (-6.0 + 12.0 * z) / (kDiskRim * kDiskRim);
    return 1.0 - 3.0 * z * z + 2.0 * z * z * z;
}

// penalty of x against primitive j (planes first, then disks centred at cc)
__device__ double penalty_dev(const ContactArgs& c, int j, const double* x, const double* cc, double* gr,
                              double* H) {
    const bool plane = j < c.nplanes;
    const double* p = plane ? c.pp[j] : cc;
    const double* nn = plane ? c.pn[j] : c.dn[j - c.nplanes];
    const double v0 = x[0] - p[0], v1 = x[1] - p[1], v2 = x[2] - p[2];
    const double d = nn[0] * v0 + nn[1] * v1 + nn[2] * v2;
    if (!(d < 0.0)) return 0.0;
    double w = 1.0, ds = 0.0, d2s = 0.0, rho = 0.0, th[3] = {0.0, 0.0, 0.0};
    if (!plane) {
        const double t0 = v0 - d * nn[0], t1 = v1 - d * nn[1], t2 = v2 - d * nn[2];
        rho = sqrt(t0 * t0 + t1 * t1 + t2 * t2);
        w = rim_weight_dev(rho, c.dr[j - c.nplanes], &ds, &d2s);
        if (w == 0.0) return 0.0;
        if (rho > 0.0) {
            th[0] = t0 / rho;
            th[1] = t1 / rho;
            th[2] = t2 / rho;
        } else {
            ds = d2s = 0.0;
        }
    }
    if (gr)
#pragma unroll
        for (int a = 0; a < 3; ++a) gr[a] += c.k * w * d * nn[a] + 0.5 * c.k * d * d * ds * th[a];
    if (H) {
        const bool rim = ds != 0.0 || d2s != 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                double h = c.k * w * nn[a] * nn[b];
                if (rim) {
                    const double P = (a == b ? 1.0 : 0.0) - nn[a] * nn[b] - th[a] * th[b];
                    h += c.k * d * ds * (nn[a] * th[b] + th[a] * nn[b]) +
                         0.5 * c.k * d * d * (d2s * th[a] * th[b] + ds * P / rho);
                }
                H[3 * a + b] += h;
            }
    }
    return 0.5 * c.k * w * d * d;
}

__device__ double friction_dev(const ContactArgs& c, int i, const double* x, const double* xprev, double dt,
                               double* gr, double* H) {
    const double N = c.lagN[i];
    if (!(N > 0.0) || !(c.mu > 0.0)) return 0.0;
    const double* nn = c.lagn + 3 * (size_t)i;
    const double* D = c.lagD + 3 * (size_t)i;
    double u[3], uT[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) u[a] = x[a] - xprev[a] - D[a];
    const double un = nn[0] * u[0] + nn[1] * u[1] + nn[2] * u[2];
#pragma unroll
    for (int a = 0; a < 3; ++a) uT[a] = u[a] - un * nn[a];
    const double y = sqrt(uT[0] * uT[0] + uT[1] * uT[1] + uT[2] * uT[2]);
    const double eps = c.eps_v * dt, muN = c.mu * N;
    double f1y, c2, f0;  // f1(y)/y, coefficient of uT uT^T, f0(y)
    if (y < eps) {
        f1y = -y / (eps * eps) + 2.0 / eps;
        c2 = y > 0.0 ? -1.0 / (y * eps * eps) : 0.0;
        f0 = -y * y * y / (3.0 * eps * eps) + y * y / eps + eps / 3.0;
    } else {
        f1y = 1.0 / y;
        c2 = -1.0 / (y * y * y);
        f0 = y;
    }
    if (gr)
#pragma unroll
        for (int a = 0; a < 3; ++a) gr[a] += muN * f1y * uT[a];
    if (H)
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b)
                H[3 * a + b] += muN * (f1y * ((a == b ? 1.0 : 0.0) - nn[a] * nn[b]) + c2 * uT[a] * uT[b]);
    return muN * f0;
}

// symmetric 3x3 with negative eigenvalues clamped to 0
__device__ void project_psd3_dev(double* H) {
    double A[3][3], V[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) A[a][b] = 0.5 * (H[3 * a + b] + H[3 * b + a]);
    jacobi3<double>(A, V);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            double h = 0.0;
            for (int k = 0; k < 3; ++k)
                if (A[k][k] > 0.0) h += A[k][k] * V[a][k] * V[b][k];
            H[3 * a + b] = h;
        }
}

// all contact terms of candidate i at position x (xprev: its qprev)
__device__ double contact_node(const ContactArgs& c, int i, const double* x, const double* xprev, double dt,
                               double* gr, double* H) {
    double e = 0.0;
    for (int j = 0; j < c.nplanes + c.ndisks; ++j)
        e += penalty_dev(c, j, x, j < c.nplanes ? nullptr : c.dc[j - c.nplanes], gr, H);
    if (dt != 0.0) e += friction_dev(c, i, x, xprev, dt, gr, H);
    return e;
}

// gradient into gcon (surface DOFs; masked DOFs 0) and, with vals, the
// Hessian into each node's self block (sliced layout: sbase/self_slot;
// scalar CSR: rowptr/cols)
template <typename HT>
__global__ __launch_bounds__(kBlock) void k_contact_terms(ContactArgs c, const double* q, const double* qprev,
                                                          double dt, double* gcon, HT* vals, const int* sbase,
                                                          const int* self_slot, const int* rowptr, const int* cols) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.nsurf) return;
    double gr[3] = {0.0, 0.0, 0.0}, H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int node = c.surf[i];
    contact_node(c, i, q + 3 * (size_t)node, qprev + 3 * (size_t)node, dt, gr, vals ? H : nullptr);
    if (vals) project_psd3_dev(H);
    bool m[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        m[a] = c.mask && c.mask[3 * node + a];
        gcon[3 * node + a] = m[a] ? 0.0 : gr[a];
    }
    if (!vals) return;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            if (m[a] || m[b] || H[3 * a + b] == 0.0) continue;
            long long idx = -1;
            if (sbase) {
                idx = block_index(sbase, node, self_slot[node], a, b);
            } else {
                const int R = 3 * node + a, C = 3 * node + b;
                int lo = rowptr[R], hi = rowptr[R + 1] - 1;
                while (lo <= hi) {
                    const int mid = (lo + hi) >> 1;
                    if (cols[mid] == C) {
                        idx = mid;
                        break;
                    }
                    if (cols[mid] < C)
                        lo = mid + 1;
                    else
                        hi = mid - 1;
                }
            }
            if (idx >= 0) vals[idx] += (HT)H[3 * a + b];
        }
}

// contact energy (diff: E(q + alpha dq) - E(q) per node, then summed)
__global__ __launch_bounds__(kBlock) void k_contact_energy(ContactArgs c, const double* q, const double* dq,
                                                           double alpha, int diff, const double* qprev, double dt,
                                                           Reduce red) {
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < c.nsurf; i += gridDim.x * blockDim.x) {
        const int node = c.surf[i];
        const double* x = q + 3 * (size_t)node;
        const double* xp = qprev + 3 * (size_t)node;
        if (diff) {
            double xt[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) xt[a] = x[a] + alpha * dq[3 * (size_t)node + a];
            v[0] += contact_node(c, i, xt, xp, dt, nullptr, nullptr) - contact_node(c, i, x, xp, dt, nullptr, nullptr);
        } else {
            v[0] += contact_node(c, i, x, xp, dt, nullptr, nullptr);
        }
    }
    block_partial<1>(v, red.partials);
}

// friction lag from qprev: the primitive with the largest penalty force
__global__ __launch_bounds__(kBlock) void k_contact_lag(ContactArgs c, const double* qprev) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.nsurf) return;
    const double* x = qprev + 3 * (size_t)c.surf[i];
    double best = 0.0, bn[3] = {0.0, 0.0, 0.0}, bD[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < c.nplanes + c.ndisks; ++j) {
        double gr[3] = {0.0, 0.0, 0.0};
        const bool plane = j < c.nplanes;
        penalty_dev(c, j, x, plane ? nullptr : c.dc_last[j - c.nplanes], gr, nullptr);
        const double* nn = plane ? c.pn[j] : c.dn[j - c.nplanes];
        const double f = fabs(gr[0] * nn[0] + gr[1] * nn[1] + gr[2] * nn[2]);
        if (f > best) {
            best = f;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                bn[a] = nn[a];
                bD[a] = plane ? 0.0 : c.dc[j - c.nplanes][a] - c.dc_last[j - c.nplanes][a];
            }
        }
    }
    c.lagN[i] = best;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        c.lagn[3 * (size_t)i + a] = bn[a];
        c.lagD[3 * (size_t)i + a] = bD[a];
    }
}

// ===========================================================================
// Block-Jacobi preconditioner: 3x3 diagonal node blocks inverted in FP64
// ===========================================================================
template <typename HT>
__global__ __launch_bounds__(kBlock) void k_block_jacobi(int N, const int* rowptr, const int* cols, const int* diag,
                                                         const HT* vals, HT* Minv) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    double B[9];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int R = 3 * v + c;
        const int di = diag[R];
        const int lo = rowptr[R], hi = rowptr[R + 1];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int k = di + (d - c);
            B[3 * c + d] = (k >= lo && k < hi && cols[k] == 3 * v + d) ? (double)vals[k] : 0.0;
        }
    }
    M3<double> A;
#pragma unroll
    for (int i = 0; i < 9; ++i) A.m[i] = B[i];
    const M3<double> C = cofactor3(A);
    const double det = A.m[0] * C.m[0] + A.m[1] * C.m[1] + A.m[2] * C.m[2];
    double out[9];
    if (det != 0.0 && isfinite(det)) {
        // A^-1 (i, j) = C(j, i) / det
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) out[3 * i + j] = C(j, i) / det;
    } else {
#pragma unroll
        for (int i = 0; i < 9; ++i) out[i] = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) out[4 * i] = (B[4 * i] != 0.0) ? 1.0 / B[4 * i] : 1.0;
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) Minv[(size_t)9 * v + i] = (HT)out[i];
}

// block-Jacobi from the sliced node-block layout: the node's self block
template <typename HT>
__global__ __launch_bounds__(kBlock) void k_block_jacobi_blk(int N, const int* sbase, const int* self_slot,
                                                             const HT* vals, HT* Minv) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const long long base = block_index(sbase, v, self_slot[v], 0, 0);
    M3<double> A;
#pragma unroll
    for (int k = 0; k < 9; ++k) A.m[k] = (double)vals[base + (long long)kSlice * k];
    const M3<double> C = cofactor3(A);
    const double det = A.m[0] * C.m[0] + A.m[1] * C.m[1] + A.m[2] * C.m[2];
    double out[9];
    if (det != 0.0 && isfinite(det)) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) out[3 * i + j] = C(j, i) / det;
    } else {
#pragma unroll
        for (int i = 0; i < 9; ++i) out[i] = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) out[4 * i] = (A.m[4 * i] != 0.0) ? 1.0 / A.m[4 * i] : 1.0;
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) Minv[(size_t)9 * v + i] = (HT)out[i];
}

// CSR values from the sliced node-block layout (export for inspection).
// Row R = 3v + c of a kept node lists 3 columns per kept neighbour in slot
// order; a Dirichlet node's row holds only (R, R).
template <typename HT>
__global__ void k_block_to_csr(int n, const int* rowptr, const int* cols, const int* sbase, const int* self_slot,
                               const int* mask, const HT* bvals, double* out) {
    const int R = blockIdx.x * blockDim.x + threadIdx.x;
    if (R >= n) return;
    const int v = R / 3, c = R % 3;
    const bool dir = mask && mask[R];
    for (int p = rowptr[R]; p < rowptr[R + 1]; ++p) {
        const int C = cols[p];
        long long idx;
        if (dir)
            idx = block_index(sbase, v, self_slot[v], c, c);
        else
            idx = block_index(sbase, v, (p - rowptr[R]) / 3, c, C % 3);
        out[p] = (double)bvals[idx];
    }
}

// ===========================================================================
// PCG, Chronopoulos-Gear form: per iteration k one fused vector pass U_k and
// one SpMV A_k.  FP64 dot products.  Scalars live on the device: U_k and A_k
// start by summing the previous kernel's per-block partials (identically in
// every block), so no kernel needs a grid-wide fence.  Ping-pong state:
//   U_k reads  st[(k-1)%2] = {gamma_{k-2}, alpha_{k-2}}, partU[(k-1)%2] =
//              gamma_{k-1}, partA = delta_{k-1}; writes st[k%2] and
//              partU[k%2] = (gamma_k, rr_k)
//   A_k reads  partU[k%2] (convergence), writes partA = delta_k
// A_0 / U_1 are the first iteration (alpha_0 = gamma_0/delta_0, beta_0 = 0).
// ===========================================================================

// ===========================================================================
// Adjoint of one backward Euler step (no reference counterpart).  With the
// step residual r(q1; q0, v0, theta) = 0 and lambda = H^-1 (qbar1 + vbar1/dt)
// (H = dr/dq1, symmetric):
//   qbar0 = -(dr/dq0)^T lambda - vbar1/dt,  vbar0 = -(dr/dv0)^T lambda,
//   thetabar = -(dr/dtheta)^T lambda.
// dr/dq0 = -(m/dt^2 + alpha m/dt), dr/dv0 = -m/dt on free DOFs; a Dirichlet
// row r_i = q1_i - q0_i - dt bc_i gives dr/dq0 = -1, dr/dv0 = 0.
// ===========================================================================
__global__ __launch_bounds__(kBlock) void k_adj_state(int n, double dt, double damping, const double* mass,
                                                      const int* mask, const double* lam, const double* vbar,
                                                      double* qbar0, double* vbar0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double m = mass[i], l = lam[i];
    const bool dir = mask && mask[i];
    qbar0[i] = (dir ? l : (m / (dt * dt) + damping * m / dt) * l) - vbar[i] / dt;
    vbar0[i] = dir ? 0.0 : m / dt * l;
}

// per-element parameter adjoints of the tet4 residual at q1:
//   Ebar_e = -lambda_e . dg_e/dE (Neo-Hookean gradient with the Lame
//   constants' derivatives dmu/dE = 1/(2(1+nu)), dlam/dE = nu/((1+nu)(1-2nu)));
//   abar_e = -lambda_e . dg_e/da (active term with a = 1: V0 sigma (F f)(f.g_k)),
//   summed per muscle group by k_group_sum in a fixed order.
// Dirichlet DOFs take no element terms (their rows are the Dirichlet rows).
__global__ __launch_bounds__(kBlock) void k_adj_params_tet(ElemArgs a, const double* nu, const double* lam,
                                                           double* Ebar, double* abar_e) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.E) return;
    int nd[4];
    double X[4][3], x[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        nd[k] = a.conn[(size_t)k * a.E + e];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            X[k][c] = a.rest[3 * nd[k] + c];
            x[k][c] = a.q[3 * nd[k] + c];
        }
    }
    M3<double> Dm;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) Dm(r, c) = X[c][r] - X[3][r];
    const M3<double> Cm = cofactor3(Dm);
    const double detm = Dm.m[0] * Cm.m[0] + Dm.m[1] * Cm.m[1] + Dm.m[2] * Cm.m[2];
    const double V0 = fabs(detm) / 6.0;
    double g[4][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) g[k][c] = Cm(c, k) / detm;
#pragma unroll
    for (int c = 0; c < 3; ++c) g[3][c] = -(g[0][c] + g[1][c] + g[2][c]);
    M3<double> D;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double acc = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) acc += (x[k][r] - X[k][r]) * g[k][c];
            D(r, c) = acc;
        }
    const double v = nu[e];
    const double mu1 = 1.0 / (2.0 * (1.0 + v)), lam1 = v / ((1.0 + v) * (1.0 - 2.0 * v));
    const NHPoint<double> p = nh_point(D);
    const M3<double> P1 = nh_stress(p, mu1, lam1);
    double fa[3] = {0.0, 0.0, 0.0}, Ff[3] = {0.0, 0.0, 0.0};
    const int gi = a.agroup ? a.agroup[e] : -1;
    if (gi >= 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) fa[c] = a.afib[(size_t)c * a.E + e];
#pragma unroll
        for (int r = 0; r < 3; ++r) Ff[r] = fa[r] + D(r, 0) * fa[0] + D(r, 1) * fa[1] + D(r, 2) * fa[2];
    }
    double sE = 0.0, sA = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double fg = fa[0] * g[k][0] + fa[1] * g[k][1] + fa[2] * g[k][2];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const int dof = 3 * nd[k] + r;
            if (!dof_kept(a, dof)) continue;
            const double l = lam[dof];
            sE += l * V0 * (P1(r, 0) * g[k][0] + P1(r, 1) * g[k][1] + P1(r, 2) * g[k][2]);
            if (gi >= 0) sA += l * V0 * a.sigma * Ff[r] * fg;
        }
    }
    Ebar[e] = -sE;
    if (abar_e) abar_e[e] = -sA;
}

// out[g] = sum of val[e] over elements of group g, in element order per
// thread and a fixed tree across the block (one block per group)
__global__ __launch_bounds__(kBlock) void k_group_sum(int E, const int* group, const double* val, double* out) {
    const int gsel = blockIdx.x;
    double s[1] = {0.0};
    for (int e = threadIdx.x; e < E; e += blockDim.x)
        if (group[e] == gsel) s[0] += val[e];
    block_sum<1>(s);
    if (threadIdx.x == 0) out[gsel] = s[0];
}


struct PcgPart {
    double* u[2];   // 2 slots each: gamma, rr
    double* a;      // 1 slot: delta
};

// One rollout's PCG data (sliced node-block matrix).
template <typename HT>
struct PcgView {
    int N, n, nslices;
    const int* sbase;
    const int* bcol;
    const short* bcol16 = nullptr;  // non-null: the SpMV reads 16-bit column offsets
    const HT* vals;
    const HT* Minv;
    HT *x, *r, *u, *p, *s, *w;
    PcgCtl* ctl;
    PcgState* st;  // st[0], st[1]
    double rtol;
};

// Where one kernel's blocks write their partial sums (slot k of block lb at
// base[k * stride + lb]) and where the sums it consumes live.
struct PartIO {
    double* out;        // this kernel's partials
    int stride;
    const double* inU;  // consumed gamma/rr partials, with their producer's block count
    int nbU;
    const double* inA;  // consumed delta partials
    int nbA;
    const double* inC = nullptr;  // coarse part of gamma (two-level preconditioner) or null
    int nbC = 0;
};

template <int NV>
__device__ __forceinline__ void block_partial_at(double (&v)[NV], double* base, int stride, int lb) {
    block_sum<NV>(v);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) base[k * stride + lb] = v[k];
}

// Projection of the residual onto the null space of per-node linear
// constraints (minimize_SQP's working set): a node nd with kmap[nd] = k >= 0
// has the symmetric 3x3 projector P[9k .. 9k+8], the others none.  With the
// block-Jacobi rows of those nodes replaced by P D^-1 P, PCG on the projected
// residual is CG on the constrained subspace (p, x stay in it).
template <typename HT>
struct PcgProj {
    const int* kmap = nullptr;
    const HT* P = nullptr;
};

// ---- bodies (lb / nb: this block's index and the block count of its rollout)
// init: x = 0, r = sign*b (projected with PROJ), p = s = 0, u = M r; partials (r.u, r.r)
template <typename HT, bool PROJ = false>
__device__ void pcg_init_body(const PcgView<HT>& v, const double* b, double sign, PartIO io, int lb, int nb,
                              PcgProj<HT> pj = {}) {
    double acc[2] = {0.0, 0.0};
    for (int nd = lb * blockDim.x + threadIdx.x; nd < v.N; nd += nb * blockDim.x) {
        HT rr[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) rr[c] = (HT)(sign * b[3 * nd + c]);
        if constexpr (PROJ) {
            const int k = pj.kmap[nd];
            if (k >= 0) {
                const HT* P = pj.P + 9 * (size_t)k;
                HT t[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) t[c] = P[3 * c] * rr[0] + P[3 * c + 1] * rr[1] + P[3 * c + 2] * rr[2];
#pragma unroll
                for (int c = 0; c < 3; ++c) rr[c] = t[c];
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            v.r[3 * nd + c] = rr[c];
            v.x[3 * nd + c] = 0;
            v.p[3 * nd + c] = 0;
            v.s[3 * nd + c] = 0;
        }
        const HT* M = v.Minv + (size_t)9 * nd;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const HT uu = M[3 * c] * rr[0] + M[3 * c + 1] * rr[1] + M[3 * c + 2] * rr[2];
            v.u[3 * nd + c] = uu;
            acc[0] += (double)rr[c] * (double)uu;
            acc[1] += (double)rr[c] * (double)rr[c];
        }
    }
    block_partial_at<2>(acc, io.out, io.stride, lb);
}

// Sums of the partial sets a kernel consumes, in one pass: t[0], t[1] the two
// slots of inU; t[2] inA (when given); t[3] inC (when given), each over its
// own producer's block count in ascending order, then one workgroup sum.
__device__ __forceinline__ void sum_partials_multi(const PartIO& io, double (&t)[4]) {
    __shared__ double bc[4];
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    const int nbA = io.inA ? io.nbA : 0, nbC = io.inC ? io.nbC : 0;
    const int nmax = max(io.nbU, max(nbA, nbC));
    const int step = blockDim.x;
    // kPartTrips slots per thread and pass: a pass issues all its loads
    // before its first add (one cache round trip per pass instead of one per
    // slot); each thread adds its slots in ascending order
    constexpr int kPartTrips = 4;
    for (int i0 = threadIdx.x; i0 < nmax; i0 += kPartTrips * step) {
        double u0[kPartTrips], u1[kPartTrips], a[kPartTrips], cc[kPartTrips];
#pragma unroll
        for (int k = 0; k < kPartTrips; ++k) {
            const int i = i0 + k * step;
            u0[k] = u1[k] = a[k] = cc[k] = 0.0;
            if (i < io.nbU) {
                u0[k] = io.inU[i];
                u1[k] = io.inU[io.stride + i];
            }
            if (i < nbA) a[k] = io.inA[i];
            if (i < nbC) cc[k] = io.inC[i];
        }
#pragma unroll
        for (int k = 0; k < kPartTrips; ++k) {
            const int i = i0 + k * step;
            if (i < io.nbU) {
                s[0] += u0[k];
                s[1] += u1[k];
            }
            if (i < nbA) s[2] += a[k];
            if (i < nbC) s[3] += cc[k];
        }
    }
    block_sum<4>(s);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 4; ++k) bc[k] = s[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) t[k] = bc[k];
    __syncthreads();
}

// ctl->done as one value for the whole workgroup: the flag can be set by
// another block of the same launch while this one runs, and the waves of a
// workgroup must take the same early exit (a persistent launch goes on to
// further virtual blocks, barriers included, after an exit)
__device__ __forceinline__ bool block_done(const PcgCtl* ctl) {
    __shared__ int flag;
    if (threadIdx.x == 0) flag = ctl->done;
    __syncthreads();
    const bool d = flag != 0;
    __syncthreads();
    return d;
}

// SpMV prologue: returns true when the solve is over (converged, breakdown,
// or a zero right-hand side).  Block 0 of the rollout records the outcome.
__device__ __forceinline__ bool spmv_prologue(PcgCtl* ctl, PcgState* st, const double* partU, int stride, int nbU,
                                              int first, double rtol, bool lead, const double* partC = nullptr,
                                              int nbC = 0) {
    if (block_done(ctl)) return true;
    double t[2];
    {
        const PartIO io{nullptr, stride, partU, nbU, nullptr, 0, partC, nbC};
        double u[4];
        sum_partials_multi(io, u);  // gamma_k, rr_k, -, coarse part of gamma_k
        t[0] = u[0] + u[3];
        t[1] = u[1];
    }
    if (first) {
        const bool zero = !(t[1] > 0.0) || !(t[0] > 0.0);
        if (lead) {
            st->rr0 = t[1];
            st->rr = t[1];
            st->tol2 = rtol * rtol * t[1];
            st->gamma = t[0];
            st->alpha = 0.0;
            st->iter = 0;
            if (zero) {
                ctl->done = 1;
                ctl->iter = 0;
                ctl->rr = t[1];
                ctl->breakdown = (t[1] > 0.0) ? 1 : 0;
            }
        }
        return zero;
    }
    const bool nan = !(t[1] == t[1]);
    if (t[1] <= st->tol2 || nan) {
        if (lead) {
            ctl->done = 1;
            ctl->iter = st->iter;
            ctl->rr = t[1];
            ctl->breakdown = nan ? 1 : 0;
        }
        return true;
    }
    return false;
}

// A_k: w = A u on the sliced node-block layout (one lane per node, one
// wavefront per slice of 64 nodes, unit-stride value and index loads)
template <typename HT>
__device__ void pcg_spmv_body(const PcgView<HT>& v, PcgState* st, PartIO io, int first, int lb, int nb) {
    const bool lead = lb == 0 && threadIdx.x == 0;
    if (spmv_prologue(v.ctl, st, io.inU, io.stride, io.nbU, first, v.rtol, lead, io.inC, io.nbC)) return;
    double acc[1] = {0.0};
    const int lane = threadIdx.x & (kSlice - 1), wv = threadIdx.x / kSlice;
    constexpr int kPer = kBlock / kSlice;  // slices per block
    const int ngroups = (v.nslices + kPer - 1) / kPer;
    const int per_xcd = (ngroups + kNumXcd - 1) / kNumXcd;
    for (int t = lb; t < kNumXcd * per_xcd; t += nb) {
        const int xcd = t % kNumXcd, local = t / kNumXcd;
        const int g = xcd * per_xcd + local;
        const int sl = g * kPer + wv;
        if (local >= per_xcd || g >= ngroups || sl >= v.nslices) continue;  // uniform per wavefront
        const int base = v.sbase[sl];
        const int width = (v.sbase[sl + 1] - base) / kSlice;
        const HT* vp = v.vals + 9LL * base + lane;
        const int node = sl * kSlice + lane;
        HT a0 = 0, a1 = 0, a2 = 0;
        auto slot = [&](int j, int nbr) {
            const HT x0 = v.u[3 * nbr], x1 = v.u[3 * nbr + 1], x2 = v.u[3 * nbr + 2];
            const HT* bb = vp + (long long)kSlice * 9 * j;
            a0 += bb[0] * x0 + bb[kSlice] * x1 + bb[2 * kSlice] * x2;
            a1 += bb[3 * kSlice] * x0 + bb[4 * kSlice] * x1 + bb[5 * kSlice] * x2;
            a2 += bb[6 * kSlice] * x0 + bb[7 * kSlice] * x1 + bb[8 * kSlice] * x2;
        };
        if (v.bcol16) {  // 2-byte column offsets: half the index traffic, the same neighbours
            const short* cp = v.bcol16 + base + lane;
            for (int j = 0; j < width; ++j) slot(j, node + cp[kSlice * j]);
        } else {
            const int* cp = v.bcol + base + lane;
            for (int j = 0; j < width; ++j) slot(j, cp[kSlice * j]);
        }
        if (node < v.N) {
            v.w[3 * node] = a0;
            v.w[3 * node + 1] = a1;
            v.w[3 * node + 2] = a2;
            acc[0] += (double)a0 * (double)v.u[3 * node] + (double)a1 * (double)v.u[3 * node + 1] +
                      (double)a2 * (double)v.u[3 * node + 2];
        }
    }
    block_partial_at<1>(acc, io.out, io.stride, lb);
}

// U_k: p = u + beta p; s = w + beta s; x += alpha p; r -= alpha s; u = M r;
// partials (r.u, r.r).  One thread per DOF (unit-stride vector traffic), the
// three residual components of a node meet in LDS for u = M r; kUnroll DOF
// groups per trip keep their loads in flight together.
#ifndef SORS_UPD_THREADS
#define SORS_UPD_THREADS 384
#endif
constexpr int kUpdBlock = SORS_UPD_THREADS;  // multiple of 3 (whole nodes)
template <typename HT, bool PROJ = false>
__device__ void pcg_update_body(const PcgView<HT>& v, PcgState* st_prev, PcgState* st_next, PartIO io, int first,
                                int lb, int nb, PcgProj<HT> pj = {}) {
#ifndef SORS_UPD_UNROLL
#define SORS_UPD_UNROLL 8
#endif
    constexpr int kUnroll = SORS_UPD_UNROLL;
    const int n = v.n;
    const int stride = nb * kUpdBlock;
    const int c = threadIdx.x % 3, first3 = threadIdx.x - c;
    // software pipeline: the first trip's vectors and preconditioner rows are
    // loaded before the prologue's partial sums, whose latency they overlap
    HT pp[kUnroll], ss[kUnroll], uu0[kUnroll], ww[kUnroll], rr0[kUnroll], xx[kUnroll], mm[kUnroll][3];
    auto load = [&](int base) {
#pragma unroll
        for (int q = 0; q < kUnroll; ++q) {
            const int i = base + q * stride + threadIdx.x;
            if (i < n) {
                uu0[q] = v.u[i];
                pp[q] = v.p[i];
                ww[q] = v.w[i];
                ss[q] = v.s[i];
                rr0[q] = v.r[i];
                xx[q] = v.x[i];
                const HT* M = v.Minv + 3 * (size_t)i;  // row c of node i/3: Minv[9 (i/3) + 3 c] = Minv[3 i]
                mm[q][0] = M[0];
                mm[q][1] = M[1];
                mm[q][2] = M[2];
            }
        }
    };
    if (block_done(v.ctl)) return;  // before the prefetch: launches after convergence cost no traffic
    int base = lb * kUpdBlock;
    if (base < n) load(base);
    double g[2], d[1];
    {
        // gamma_{k-1} and rr_{k-1} (inU), delta_{k-1} (inA), coarse part of
        // gamma_{k-1} (inC): one pass over the three partial sets and one
        // workgroup sum; each slot keeps its own fixed summation order
        double t[4];
        sum_partials_multi(io, t);
        g[0] = t[0] + t[3];
        g[1] = t[1];
        d[0] = t[2];
    }
    double alpha, beta;
    bool bad;
    if (first) {
        beta = 0.0;
        alpha = g[0] / d[0];
        bad = !(d[0] > 0.0);
    } else {
        beta = g[0] / st_prev->gamma;
        const double den = d[0] - beta * g[0] / st_prev->alpha;
        alpha = g[0] / den;
        bad = !(den > 0.0) || !(g[0] > 0.0);
    }
    const bool lead = lb == 0 && threadIdx.x == 0;
    if (bad) {
        if (lead) {
            v.ctl->done = 1;
            v.ctl->breakdown = 1;
            v.ctl->iter = st_prev->iter;
            v.ctl->rr = g[1];
        }
        return;
    }
    if (lead) {
        st_next->gamma = g[0];
        st_next->alpha = alpha;
        st_next->rr0 = st_prev->rr0;
        st_next->tol2 = st_prev->tol2;
        st_next->rr = g[1];
        st_next->iter = st_prev->iter + 1;
    }
    const HT a = (HT)alpha, bt = (HT)beta;
    __shared__ HT rs[kUnroll][kUpdBlock];
    double acc[2] = {0.0, 0.0};
    for (; base < n; base += kUnroll * stride) {
#pragma unroll
        for (int q = 0; q < kUnroll; ++q) {
            const int i = base + q * stride + threadIdx.x;
            HT rr = 0;
            if (i < n) {
                const HT pn = uu0[q] + bt * pp[q];
                const HT sn = ww[q] + bt * ss[q];
                v.p[i] = pn;
                v.s[i] = sn;
                v.x[i] = xx[q] + a * pn;
                rr = rr0[q] - a * sn;
                if constexpr (!PROJ) v.r[i] = rr;
            }
            rs[q][threadIdx.x] = rr;
        }
        __syncthreads();
        if constexpr (PROJ) {
            // r <- P r on constrained nodes (the node's three residual
            // components meet in LDS as for u = M r)
            HT rp[kUnroll];
#pragma unroll
            for (int q = 0; q < kUnroll; ++q) {
                const int i = base + q * stride + threadIdx.x;
                rp[q] = rs[q][threadIdx.x];
                if (i < n) {
                    const int k = pj.kmap[i / 3];
                    if (k >= 0) {
                        const HT* P = pj.P + 9 * (size_t)k + 3 * c;
                        rp[q] = P[0] * rs[q][first3] + P[1] * rs[q][first3 + 1] + P[2] * rs[q][first3 + 2];
                    }
                    v.r[i] = rp[q];
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < kUnroll; ++q) rs[q][threadIdx.x] = rp[q];
            __syncthreads();
        }
#pragma unroll
        for (int q = 0; q < kUnroll; ++q) {
            const int i = base + q * stride + threadIdx.x;
            if (i < n) {
                const HT rr = rs[q][threadIdx.x];
                const HT un = mm[q][0] * rs[q][first3] + mm[q][1] * rs[q][first3 + 1] + mm[q][2] * rs[q][first3 + 2];
                v.u[i] = un;
                acc[0] += (double)rr * (double)un;
                acc[1] += (double)rr * (double)rr;
            }
        }
        __syncthreads();
        const int next = base + kUnroll * stride;
        if (next < n) load(next);
    }
    block_partial_at<2>(acc, io.out, io.stride, lb);
}

// ===========================================================================
// Two-level preconditioner: M^-1 = D^-1 + Z E^-1 Z^T with Z the rigid-body
// modes of spatial node aggregates (coarse.h) and E = Z^T A Z dense.
// Z_i [t; w] = t + w x d_i,  Z_i^T r_i = [r_i; d_i x r_i],  d_i = (x_i - c_I)/s_I
// ===========================================================================
struct CoarseArgs {
    int nagg, nc;
    const float4* dn;     // [N] (d_i, aggregate bits); d_i = 0 where rotations are dropped
    const int* aptr;      // aggregate -> anodes
    const int* anodes;    // nodes grouped by aggregate
    double* rc;           // [nc] Z^T r
};

__device__ __forceinline__ int coarse_agg(float4 d) { return __float_as_int(d.w); }

// Galerkin product by chunks: chunk g of pair (I, J) sums Z_i^T A_ij Z_j over
// at most 256 of the pair's fine blocks, one wavefront per chunk with a
// fixed-order wavefront reduction; k_coarse_gather adds a pair's chunks in
// order (deterministic).
template <typename HT>
__global__ __launch_bounds__(kBlock) void k_coarse_galerkin(int nchunks, const int* gptr, const int* ci,
                                                            const int* cj, const int* cbase, const float4* dn,
                                                            const HT* vals, double* gpart) {
    const int g = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (g >= nchunks) return;  // whole wavefronts
    double acc[36];
#pragma unroll
    for (int k = 0; k < 36; ++k) acc[k] = 0.0;
    for (int t = gptr[g] + lane; t < gptr[g + 1]; t += 64) {
        const long long base = cbase[t];
        double A[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) A[a][b] = (double)vals[base + (long long)kSlice * (3 * a + b)];
        const float4 fi = dn[ci[t]], fj = dn[cj[t]];
        const double di[3] = {fi.x, fi.y, fi.z}, dj[3] = {fj.x, fj.y, fj.z};
        // T = A Z_j (3 x 6): columns 0-2 = A, column 3+k = A (e_k x d_j)
        double T[3][6];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            T[r][0] = A[r][0];
            T[r][1] = A[r][1];
            T[r][2] = A[r][2];
            // e_x x d = (0, -d_z, d_y), e_y x d = (d_z, 0, -d_x), e_z x d = (-d_y, d_x, 0)
            T[r][3] = -A[r][1] * dj[2] + A[r][2] * dj[1];
            T[r][4] = A[r][0] * dj[2] - A[r][2] * dj[0];
            T[r][5] = -A[r][0] * dj[1] + A[r][1] * dj[0];
        }
        // Z_i^T T: rows 0-2 = T, row 3+k = (e_k x d_i) . T[:, col]
#pragma unroll
        for (int col = 0; col < 6; ++col) {
            acc[0 * 6 + col] += T[0][col];
            acc[1 * 6 + col] += T[1][col];
            acc[2 * 6 + col] += T[2][col];
            acc[3 * 6 + col] += -di[2] * T[1][col] + di[1] * T[2][col];
            acc[4 * 6 + col] += di[2] * T[0][col] - di[0] * T[2][col];
            acc[5 * 6 + col] += -di[1] * T[0][col] + di[0] * T[1][col];
        }
    }
#pragma unroll
    for (int k = 0; k < 36; ++k) {
        double x = acc[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        acc[k] = x;
    }
    if (lane < 36) {
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < 36; ++k)
            if (k == lane) v = acc[k];
        gpart[36 * (size_t)g + lane] = v;
    }
}

// E(I, J) block entries: one thread per (pair, entry), chunks in order.
// Unknowns of aggregates without rotations get a unit diagonal (decoupled:
// their Z columns are zero).
__global__ void k_coarse_gather(int npairs, int nc, const int* pairI, const int* pairJ, const int* pchunk,
                                const int* rot, const double* gpart, double* E) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 36LL * npairs) return;
    const int p = (int)(t / 36), k = (int)(t % 36);
    const int I = pairI[p], J = pairJ[p], r = k / 6, col = k % 6;
    double v = 0.0;
    for (int g = pchunk[p]; g < pchunk[p + 1]; ++g) v += gpart[36 * (size_t)g + k];
    if (I == J && r == col && r >= 3 && !rot[I]) v = 1.0;
    E[(size_t)(6 * I + r) * nc + 6 * J + col] = v;
}

// ---- coarse inverse: blocked Gauss-Jordan on the SPD matrix E -------------
// E (nc x nc) is inverted in place of a factorisation library: blocked
// Gauss-Jordan elimination without pivoting (E is SPD, so every pivot block,
// a Schur complement of E, is SPD too), one kernel per pivot block of
// kGjPivot rows on the whole padded matrix, ping-ponging between two
// buffers.  Every output value is formed by one thread in a fixed order, so
// the inverse is the same bit for bit whatever else runs on the card.
// Cost: 2 npad^3 FP64 flops and npad / kGjPivot launches.
constexpr int kGjTile = 64;     // output tile of one workgroup (npad is a multiple)
constexpr int kGjPivot = 32;    // pivot block
constexpr int kGjThreads = 256; // 16 x 16 threads, 4 x 4 outputs each
static_assert(kGjTile % kGjPivot == 0, "pivot blocks inside output tiles");

// X = E padded to npad with the identity, E read from its lower triangle (the
// matrix a Cholesky factorisation of E would see)
__global__ void k_coarse_pad(int nc, int npad, const double* __restrict__ E, double* __restrict__ X) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)npad * npad) return;
    const int i = (int)(t / npad), j = (int)(t % npad);
    double v = (i == j) ? 1.0 : 0.0;
    if (i < nc && j < nc) v = i >= j ? E[(size_t)i * nc + j] : E[(size_t)j * nc + i];
    X[t] = v;
}

// In-place Gauss-Jordan inverse of the SPD 32 x 32 block held in P[0] (two
// LDS copies, so each elimination step costs one barrier; no pivoting, P is
// a Schur complement of E).  The inverse ends in P[kGjPivot & 1]; returns
// false on a non-positive pivot.
__device__ bool gj_invert32(double (*P)[kGjPivot][kGjPivot + 1]) {
    constexpr int NB = kGjPivot, kPer = NB * NB / kGjThreads;
    const int tid = threadIdx.x;
    bool bad = false;
    for (int t = 0; t < NB; ++t) {
        const int cur = t & 1;
        const double piv = P[cur][t][t];
        bad = bad || !(piv > 0.0);
        const double rp = 1.0 / piv;
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int e = tid + q * kGjThreads, r = e / NB, c = e % NB;
            const double prt = P[cur][r][t], ptc = P[cur][t][c];
            double v;
            if (r == t)
                v = c == t ? rp : ptc * rp;
            else
                v = c == t ? -prt * rp : P[cur][r][c] - (prt * ptc) * rp;
            P[cur ^ 1][r][c] = v;
        }
        __syncthreads();
    }
    return !bad;
}

// P^-1 of the first pivot block in[0:32, 0:32] (the later ones come from
// k_spd_gj_step).  A non-positive pivot sets *info.
__global__ __launch_bounds__(kGjThreads) void k_gj_pivot(int npad, const double* __restrict__ in,
                                                         double* __restrict__ pinv, int* info) {
    constexpr int NB = kGjPivot;
    __shared__ double P[2][NB][NB + 1];
    const int tid = threadIdx.x;
    for (int e = tid; e < NB * NB; e += kGjThreads) P[0][e / NB][e % NB] = in[(size_t)(e / NB) * npad + e % NB];
    __syncthreads();
    if (!gj_invert32(P) && tid == 0) *info = 1;
    for (int e = tid; e < NB * NB; e += kGjThreads) pinv[e] = P[NB & 1][e / NB][e % NB];
}

// One elimination step with pivot rows / columns K = [k0, k0 + kGjPivot),
// given P^-1 (pinv):
//   out[K, K] = P^-1
//   out[K, j] = P^-1 in[K, j]             (j not in K)
//   out[i, K] = -in[i, K] P^-1            (i not in K)
//   out[i, j] = in[i, j] - in[i, K] P^-1 in[K, j]
// Workgroups 1.. form one 64 x 64 output tile each.  Workgroup 0 forms the
// next pivot block K' = K + 32 of out by the same operations in the same
// order as its tile's workgroup (so the two agree bit for bit), inverts it
// and writes pinv_next, beside the tiles instead of after them.  A
// non-positive pivot sets *info.
__global__ __launch_bounds__(kGjThreads) void k_spd_gj_step(int npad, int k0, const double* __restrict__ in,
                                                            const double* __restrict__ pinv,
                                                            double* __restrict__ pinv_next,
                                                            double* __restrict__ out, int* info) {
    constexpr int NB = kGjPivot, TB = kGjTile;
    __shared__ double P[NB][NB + 1];   // P^-1
    __shared__ double B[NB][TB + 1];   // in[K, J], then P^-1 in[K, J] (P^-1 on the K columns)
    __shared__ double A[TB][NB + 1];   // in[I, K]
    const int tid = threadIdx.x;
    for (int e = tid; e < NB * NB; e += kGjThreads) P[e / NB][e % NB] = pinv[e];
    if (blockIdx.x == 0) {
        const int k1 = k0 + NB;
        if (k1 >= npad) return;
        // in[K, K'] and in[K', K]; then B1 = P^-1 in[K, K'] column by column
        double (*S)[NB][NB + 1] = reinterpret_cast<double (*)[NB][NB + 1]>(&A[0][0]);  // two 32 x 33 copies
        for (int e = tid; e < NB * NB; e += kGjThreads) {
            B[e / NB][e % NB] = in[(size_t)(k0 + e / NB) * npad + k1 + e % NB];
            S[1][e / NB][e % NB] = in[(size_t)(k1 + e / NB) * npad + k0 + e % NB];
        }
        __syncthreads();
        constexpr int kPer = NB * NB / kGjThreads;
        const int c = tid % NB, r0 = tid / NB;  // rows r0, r0 + 8, ... of column c
        double v[kPer];
        {
            double col[NB];
#pragma unroll
            for (int u = 0; u < NB; ++u) col[u] = B[u][c];
#pragma unroll
            for (int q = 0; q < kPer; ++q) {
                const int r = r0 + q * (kGjThreads / NB);
                double acc = 0.0;
#pragma unroll
                for (int u = 0; u < NB; ++u) acc += P[r][u] * col[u];
                v[q] = acc;
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kPer; ++q) B[r0 + q * (kGjThreads / NB)][c] = v[q];
        __syncthreads();
        // S = in[K', K'] - in[K', K] B1, each entry's sum over t ascending
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int r = r0 + q * (kGjThreads / NB);
            double acc = 0.0;
            for (int t = 0; t < NB; ++t) acc += S[1][r][t] * B[t][c];
            v[q] = in[(size_t)(k1 + r) * npad + k1 + c] - acc;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kPer; ++q) S[0][r0 + q * (kGjThreads / NB)][c] = v[q];
        __syncthreads();
        if (!gj_invert32(S) && tid == 0) *info = 1;
        for (int e = tid; e < NB * NB; e += kGjThreads) pinv_next[e] = S[NB & 1][e / NB][e % NB];
        return;
    }
    const int tiles = npad / TB;
    const int tile = blockIdx.x - 1;
    const int ti = tile / tiles, tj = tile % tiles;
    const int i0 = ti * TB, j0 = tj * TB;
    for (int e = tid; e < NB * TB; e += kGjThreads) B[e / TB][e % TB] = in[(size_t)(k0 + e / TB) * npad + j0 + e % TB];
    for (int e = tid; e < TB * NB; e += kGjThreads) A[e / NB][e % NB] = in[(size_t)(i0 + e / NB) * npad + k0 + e % NB];
    __syncthreads();
    // B = P^-1 in[K, J]; on the columns of K, P^-1 itself.  A thread's eight
    // outputs share one column: that column of in[K, J] is read once
    {
        constexpr int kPer = NB * TB / kGjThreads;  // rows r0, r0 + 4, ... of one column c
        const int c = tid % TB, r0 = tid / TB;
        const int j = j0 + c;
        const bool colK = j >= k0 && j < k0 + NB;
        double col[NB];
#pragma unroll
        for (int u = 0; u < NB; ++u) col[u] = B[u][c];
        double v[kPer];
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int r = r0 + q * (kGjThreads / TB);
            if (colK) {
                v[q] = P[r][j - k0];
            } else {
                double acc = 0.0;
#pragma unroll
                for (int u = 0; u < NB; ++u) acc += P[r][u] * col[u];
                v[q] = acc;
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kPer; ++q) B[r0 + q * (kGjThreads / TB)][c] = v[q];
        __syncthreads();
    }
    // the tile: rows of K take B; the others in[i, j] (0 on the K columns) - A B.
    // The 16 values of in[i, j] are loaded before the product, all in flight at once
    const int tx = tid % 16, ty = tid / 16;
    double base[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int i = i0 + ty + 16 * a, j = j0 + tx + 16 * b;
            const bool skip = (i >= k0 && i < k0 + NB) || (j >= k0 && j < k0 + NB);
            base[a][b] = skip ? 0.0 : in[(size_t)i * npad + j];
        }
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
#pragma unroll 4
    for (int t = 0; t < NB; ++t) {
        double av[4], bv[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) av[a] = A[ty + 16 * a][t];
#pragma unroll
        for (int b = 0; b < 4; ++b) bv[b] = B[t][tx + 16 * b];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] += av[a] * bv[b];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int r = ty + 16 * a, i = i0 + r;
        const bool rowK = i >= k0 && i < k0 + NB;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int c = tx + 16 * b, j = j0 + c;
            double v;
            if (rowK)
                v = B[i - k0][c];
            else
                v = base[a][b] - acc[a][b];
            out[(size_t)i * npad + j] = v;
        }
    }
}

// E^-1 = (X + X^T)/2 from X = E^-1 (ld: X's row stride), symmetric up to
// rounding: an exactly symmetric inverse, so Z E^-1 Z^T is a symmetric
// preconditioner.  The coarse solve stays in FP64 at every precision: E is
// as badly conditioned as the coarse problem, and an FP32 copy of its
// inverse need not be positive definite.
__global__ void k_coarse_symmetrize(int nc, int ld, const double* X, double* Einv, float* Einv32) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nc * nc) return;
    const int a = (int)(t / nc), b = (int)(t % nc);
    const double e = 0.5 * (X[(size_t)a * ld + b] + X[(size_t)b * ld + a]);
    Einv[t] = e;
    Einv32[t] = (float)e;  // the copy the mixed-precision apply reads
}

// rc = Z^T r: one workgroup per aggregate, its nodes in a fixed order, a
// fixed-order workgroup sum (deterministic)
#ifndef SORS_RESTRICT_THREADS
#define SORS_RESTRICT_THREADS 512
#endif
constexpr int kRestrictThreads = SORS_RESTRICT_THREADS;
template <typename HT>
__device__ void coarse_restrict_body(const CoarseArgs& c, const PcgCtl* ctl, const HT* r, int I) {
    if (ctl && block_done(ctl)) return;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    const int lo = c.aptr[I], hi = c.aptr[I + 1];
    for (int t = lo + threadIdx.x; t < hi; t += kRestrictThreads) {
        const int i = c.anodes[t];
        const float4 f = c.dn[i];
        const double d0 = f.x, d1 = f.y, d2 = f.z;
        const double r0 = (double)r[3 * i], r1 = (double)r[3 * i + 1], r2 = (double)r[3 * i + 2];
        acc[0] += r0;
        acc[1] += r1;
        acc[2] += r2;
        // (e_k x d) . r = e_k . (d x r)
        acc[3] += d1 * r2 - d2 * r1;
        acc[4] += d2 * r0 - d0 * r2;
        acc[5] += d0 * r1 - d1 * r0;
    }
    block_sum<6>(acc);  // block totals in thread 0
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 6; ++k) c.rc[6 * (size_t)I + k] = acc[k];
}

template <typename HT>
__global__ __launch_bounds__(kRestrictThreads) void k_coarse_restrict(CoarseArgs c, const PcgCtl* ctl,
                                                                      const HT* r) {
    coarse_restrict_body(c, ctl, r, blockIdx.x);
}

// xc_I = rows 6I..6I+5 of E^-1 times rc, then u += Z_I xc_I over the nodes of
// aggregate I: one workgroup per aggregate, so E^-1 is still read once per
// application and the prolongation needs no second launch.  One partial of
// gamma_c = rc . xc per workgroup; fixed-order sums throughout.  512 threads
// (restriction and apply alike): 256 and 1024 were both slower on the
// 1.92 M-tet beam (256 aggregates).
#ifndef SORS_MAX_AGGREGATES
#define SORS_MAX_AGGREGATES 512
#endif
constexpr int kMaxCoarse = 6 * SORS_MAX_AGGREGATES;  // coarse unknowns: nc^2 FP64 values are read per PCG iteration
#ifndef SORS_APPLY_THREADS
#define SORS_APPLY_THREADS 512
#endif
constexpr int kApplyThreads = SORS_APPLY_THREADS;
static_assert(kMaxCoarse / 6 <= kMaxRedBlocks, "one gamma partial per aggregate");
template <typename HT>
__device__ void coarse_apply_body(const CoarseArgs& c, const PcgCtl* ctl, const HT* __restrict__ Einv, HT* u,
                                  double* part, int I) {
    if (ctl && block_done(ctl)) return;
    __shared__ double xs[6];
    const int nc = c.nc;
    const double* __restrict__ rc = c.rc;
    const HT* __restrict__ E0 = Einv + (size_t)6 * I * nc;
    double a[6] = {0, 0, 0, 0, 0, 0};
    for (int col = threadIdx.x; col < nc; col += kApplyThreads) {
        double e[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) e[k] = (double)E0[(size_t)k * nc + col];
        const double x = rc[col];
#pragma unroll
        for (int k = 0; k < 6; ++k) a[k] += e[k] * x;
    }
    block_sum<6>(a);  // block totals in thread 0
    if (threadIdx.x == 0) {
        double g = 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            xs[k] = a[k];
            g += a[k] * rc[6 * I + k];
        }
        part[I] = g;
    }
    __syncthreads();
    const double t0 = xs[0], t1 = xs[1], t2 = xs[2], w0 = xs[3], w1 = xs[4], w2 = xs[5];
    for (int t = c.aptr[I] + threadIdx.x; t < c.aptr[I + 1]; t += kApplyThreads) {
        const int i = c.anodes[t];
        const float4 f = c.dn[i];
        const double d0 = f.x, d1 = f.y, d2 = f.z;
        u[3 * i] += (HT)(t0 + w1 * d2 - w2 * d1);
        u[3 * i + 1] += (HT)(t1 + w2 * d0 - w0 * d2);
        u[3 * i + 2] += (HT)(t2 + w0 * d1 - w1 * d0);
    }
}

template <typename HT>
__global__ __launch_bounds__(kApplyThreads) void k_coarse_apply(CoarseArgs c, const PcgCtl* ctl,
                                                                const HT* __restrict__ Einv, HT* u,
                                                                double* part) {
    coarse_apply_body(c, ctl, Einv, u, part, blockIdx.x);
}

// ---- single-rollout kernels ----------------------------------------------
template <typename HT>
__global__ __launch_bounds__(kBlock) void k_pcg_init(PcgView<HT> v, const double* b, double sign, PartIO io) {
    pcg_init_body(v, b, sign, io, blockIdx.x, gridDim.x);
}

template <typename HT>
__global__ __launch_bounds__(kBlock) void k_pcg_spmv_blk(PcgView<HT> v, PcgState* st, PartIO io, int first) {
    pcg_spmv_body(v, st, io, first, blockIdx.x, gridDim.x);
}

template <typename HT>
__global__ __launch_bounds__(kUpdBlock) void k_pcg_update(PcgView<HT> v, PcgState* st_prev, PcgState* st_next,
                                                          PartIO io, int first) {
    pcg_update_body(v, st_prev, st_next, io, first, blockIdx.x, gridDim.x);
}

// the same two kernels with the residual projected onto the constrained
// subspace (minimize_SQP)
template <typename HT>
__global__ __launch_bounds__(kBlock) void k_pcg_init_proj(PcgView<HT> v, const double* b, double sign, PartIO io,
                                                          PcgProj<HT> pj) {
    pcg_init_body<HT, true>(v, b, sign, io, blockIdx.x, gridDim.x, pj);
}

template <typename HT>
__global__ __launch_bounds__(kUpdBlock) void k_pcg_update_proj(PcgView<HT> v, PcgState* st_prev, PcgState* st_next,
                                                               PartIO io, int first, PcgProj<HT> pj) {
    pcg_update_body<HT, true>(v, st_prev, st_next, io, first, blockIdx.x, gridDim.x, pj);
}

// w = A u on scalar CSR rows (general masks), LPR lanes per row
template <typename HT, int LPR>
__global__ __launch_bounds__(kBlock) void k_pcg_spmv(int n, const int* rowptr, const int* cols, const HT* vals,
                                                     const HT* u, HT* w, PcgCtl* ctl, PcgState* st, PartIO io,
                                                     int first, double rtol) {
    if (spmv_prologue(ctl, st, io.inU, io.stride, io.nbU, first, rtol, blockIdx.x == 0 && threadIdx.x == 0, io.inC,
                      io.nbC))
        return;
    double v[1] = {0.0};
    const int sub = threadIdx.x % LPR;
    const int rows_per_block = blockDim.x / LPR;
    const int nblk_rows = (n + rows_per_block - 1) / rows_per_block;
    const int per_xcd = (nblk_rows + kNumXcd - 1) / kNumXcd;
    for (int t = blockIdx.x; t < kNumXcd * per_xcd; t += gridDim.x) {
        const int xcd = t % kNumXcd, local = t / kNumXcd;
        const int rb = xcd * per_xcd + local;
        if (rb >= nblk_rows || local >= per_xcd) continue;
        const int row = rb * rows_per_block + threadIdx.x / LPR;
        HT acc = 0;
        if (row < n) {
            const int lo = rowptr[row], hi = rowptr[row + 1];
            for (int k = lo + sub; k < hi; k += LPR) acc += vals[k] * u[cols[k]];
        }
#pragma unroll
        for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LPR);
        if (sub == 0 && row < n) {
            w[row] = acc;
            v[0] += (double)acc * (double)u[row];
        }
    }
    block_partial_at<1>(v, io.out, io.stride, blockIdx.x);
}

// ---- batched kernels: one launch over many rollouts ---------------------
// A rollout keeps its own partial buffer (gamma/rr pair for each parity,
// then delta; stride kMaxRedBlocks) and its own block count per kernel, so
// its sums are formed in the same order alone and inside a batch.
template <typename HT>
struct BMember {
    PcgView<HT> v;
    const double* rhs;  // b of A x = -b
    double* part;
    int nbI, nbA, nbU, nbC;  // nbC: aggregates of the coarse correction, 0 without one
    CoarseArgs ca;
    const HT* Einv;  // E^-1 in the rollout's precision (FP32 copy for mixed)
    double* cpart;           // gamma_c partials, two parity slots
};

// Packed launches: a batched kernel walks the virtual blocks of all rollouts
// in the launch, vb = blockIdx.x, blockIdx.x + gridDim.x, ...; the j-th
// rollout owns virtual blocks [tab[j], tab[j+1]), and the number of rollouts
// is read from device memory (*cntp).  The grid is one workgroup per virtual
// block up to a persistent cap (below).  A virtual block computes exactly
// what block lb = vb - tab[j] of the rollout's own launch computes (same
// partial slot, same summation order).  With the caps multiples of 8, virtual
// block vb runs on XCD vb % 8 as it would in a grid of its own.
constexpr int kPersistBlocks = 2048;        // 256-thread kernels: 8 per CU
constexpr int kPersistUpdBlocks = 1536;     // 384-thread update: 6 per CU
constexpr int kPersistCoarseBlocks = 1024;  // 512-thread coarse kernels: 4 per CU
static_assert(kPersistBlocks % kNumXcd == 0 && kPersistUpdBlocks % kNumXcd == 0 &&
                  kPersistCoarseBlocks % kNumXcd == 0,
              "XCD-aligned persistent grids");

__device__ __forceinline__ int find_member(const int* tab, int cnt, int b) {
    int lo = 0, hi = cnt - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// calls body(j, lb, nb) for every virtual block of this launch owned by this
// workgroup; the trailing barrier keeps one virtual block's LDS use apart
// from the next one's
template <typename F>
__device__ __forceinline__ void for_virtual_blocks(const int* tab, const int* cntp, F&& body) {
    const int cnt = *cntp;
    const int total = cnt > 0 ? tab[cnt] : 0;
    for (int vb = blockIdx.x; vb < total; vb += gridDim.x) {
        const int j = find_member(tab, cnt, vb);
        body(j, vb - tab[j], tab[j + 1] - tab[j]);
        __syncthreads();
    }
}

template <typename HT>
__global__ __launch_bounds__(kBlock) void k_b_pcg_init(const BMember<HT>* mem, const int* idx, const int* tab,
                                                       const int* cntp) {
    for_virtual_blocks(tab, cntp, [&](int j, int lb, int nb) {
        const BMember<HT>& m = mem[idx[j]];
        if (lb == 0 && threadIdx.x == 0) {
            m.v.ctl->done = 0;
            m.v.ctl->breakdown = 0;
            m.v.ctl->iter = 0;
            m.v.ctl->rr = 0.0;
        }
        const PartIO io{m.part, kMaxRedBlocks, nullptr, 0, nullptr, 0};
        pcg_init_body(m.v, m.rhs, -1.0, io, lb, nb);
    });
}

// par: parity of the iteration; first: A_0 (consumes the init partials)
template <typename HT>
__global__ __launch_bounds__(kBlock) void k_b_pcg_spmv(const BMember<HT>* mem, const int* idx, const int* tab,
                                                       const int* cntp, int par, int first) {
    for_virtual_blocks(tab, cntp, [&](int j, int lb, int nb) {
        const BMember<HT>& m = mem[idx[j]];
        const PartIO io{m.part + 4 * kMaxRedBlocks, kMaxRedBlocks, m.part + 2 * par * kMaxRedBlocks,
                        first ? m.nbI : m.nbU, nullptr, 0, m.nbC ? m.cpart + par * kMaxRedBlocks : nullptr, m.nbC};
        pcg_spmv_body(m.v, m.v.st + par, io, first, lb, nb);
    });
}

// coarse correction of the rollouts that have one (their aggregates are the virtual blocks)
template <typename HT>
__global__ __launch_bounds__(kRestrictThreads) void k_b_coarse_restrict(const BMember<HT>* mem, const int* idx,
                                                                        const int* tab, const int* cntp) {
    for_virtual_blocks(tab, cntp, [&](int j, int lb, int) {
        const BMember<HT>& m = mem[idx[j]];
        coarse_restrict_body(m.ca, m.v.ctl, (const HT*)m.v.r, lb);
    });
}

template <typename HT>
__global__ __launch_bounds__(kApplyThreads) void k_b_coarse_apply(const BMember<HT>* mem, const int* idx,
                                                                  const int* tab, const int* cntp, int par) {
    for_virtual_blocks(tab, cntp, [&](int j, int lb, int) {
        const BMember<HT>& m = mem[idx[j]];
        coarse_apply_body(m.ca, m.v.ctl, m.Einv, m.v.u, m.cpart + par * kMaxRedBlocks, lb);
    });
}

template <typename HT>
__global__ __launch_bounds__(kUpdBlock) void k_b_pcg_update(const BMember<HT>* mem, const int* idx, const int* tab,
                                                            const int* cntp, int par, int first) {
    for_virtual_blocks(tab, cntp, [&](int j, int lb, int nb) {
        const BMember<HT>& m = mem[idx[j]];
        const PartIO io{m.part + 2 * par * kMaxRedBlocks, kMaxRedBlocks, m.part + 2 * (par ^ 1) * kMaxRedBlocks,
                        first ? m.nbI : m.nbU, m.part + 4 * kMaxRedBlocks, m.nbA,
                        m.nbC ? m.cpart + (par ^ 1) * kMaxRedBlocks : nullptr, m.nbC};
        pcg_update_body(m.v, m.v.st + (par ^ 1), m.v.st + par, io, first, lb, nb);
    });
}

// ===========================================================================
// Newton vector kernels
// ===========================================================================
// ||dx||^2 and g.dx (FP64)
template <typename HT>
__global__ __launch_bounds__(kBlock) void k_step_dots(int n, const HT* dx, const double* g, double* out_dd,
                                                      double* out_gd, Reduce red) {
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double d = (double)dx[i];
        v[0] += d * d;
        v[1] += g[i] * d;
    }
    block_partial<2>(v, red.partials);
}

template <typename HT>
__global__ void k_to_f64(long long n, const HT* a, double* out) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = (double)a[i];
}

// Newton initial guess q = qPrev + h vPrev (Dirichlet DOFs: qPrev + h value)
// Newton's starting point: q = qprev + h v with v the last velocity (vold
// null) or extrapolated with the last acceleration, v + (h/h_old)(v - vold);
// Dirichlet DOFs move with their prescribed velocity
__global__ void k_predict(int n, double h, const double* qprev, const double* vprev, const double* vold,
                          double ratio, const int* mask, const double* bcval, double* q) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double v = vprev[i];
        if (vold) v += ratio * (vprev[i] - vold[i]);
        q[i] = qprev[i] + h * ((mask && mask[i]) ? bcval[i] : v);
    }
}

__global__ void k_axpy(int n, double alpha, const double* dx, double* q) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) q[i] += alpha * dx[i];
}

// TimeIntegrator::update_state (integrator.h:79-97)
__global__ void k_update_state(int n, int scheme, double dt, const double* q, double* qprev, double* vprev,
                               double* gprev, const double* gtmp, double* vold) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double v = (scheme == SORS_BACKWARD_EULER) ? (q[i] - qprev[i]) / dt : 2 * (q[i] - qprev[i]) / dt - vprev[i];
        if (vold) vold[i] = vprev[i];
        vprev[i] = v;
        qprev[i] = q[i];
        gprev[i] = gtmp[i];
    }
}

// ===========================================================================
// Static scratch of the allocation-free energy entry point
// (sors_energy_noalloc): the reduction partials of its three energy
// kernels, their three sums and the inverted-element flag.  Each device
// holds one copy from library load to process exit.  Host code never
// passes `scratch` itself to a kernel (on the host the name is a stand-in,
// not a device address): it uses the address hipGetSymbolAddress returns
// for the current device.
// ===========================================================================
constexpr int kScratchPartials = kMaxRedBlocks;       // one partial per reduction block
constexpr int kScratchSums = kScratchPartials;        // [3] element, node, contact sums
constexpr int kScratchFlag = kScratchPartials + 4;    // int flag (bit 1: inverted element)
constexpr int kScratchDoubles = kScratchPartials + 8; // 16,448 bytes per device
static __device__ __align__(256) double scratch[kScratchDoubles];

// E = element + node (+ contact) sums, in the order the host sums them for
// sors_energy
__global__ void k_energy_total(const double* s3, int with_aux, double* out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = s3[0] + s3[1] + (with_aux ? s3[2] : 0.0);
}

// plain copy (calibration of the HBM byte counters)
__global__ void k_copy(long long n4, const int4* __restrict__ src, int4* __restrict__ dst) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace sors

// ===========================================================================
// Host side: context and C ABI
// ===========================================================================
using namespace sors;

namespace {

thread_local std::string g_create_error;

// Device address of the static `scratch` array, looked up once per device
// (with that device current) and kept for the life of the process.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_addr[kMaxDevices] = {};

// address of this device's scratch copy, or null with *err set
double* scratch_address(int dev, std::string* err) {
    if (dev < 0 || dev >= kMaxDevices) {
        *err = "device ordinal " + std::to_string(dev) + " beyond the scratch table";
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(sors::scratch));
        if (e != hipSuccess || !p) {
            *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return nullptr;
        }
        g_scratch_addr[dev] = (double*)p;
    }
    return g_scratch_addr[dev];
}

struct DevMem {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevMem() {
        if (p) (void)hipFree(p);
    }
    DevMem() = default;
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    bool alloc(size_t b) {
        if (p && bytes >= b) return true;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        if (b == 0) b = 8;
        if (hipMalloc(&p, b) != hipSuccess) {
            p = nullptr;
            return false;
        }
        bytes = b;
        return true;
    }
    template <typename T>
    T* as() const {
        return (T*)p;
    }
};

struct HipError {
    std::string msg;
};

#define HIPX(expr)                                                                                      \
    do {                                                                                                \
        hipError_t _e = (expr);                                                                         \
        if (_e != hipSuccess) throw HipError{std::string(#expr) + ": " + hipGetErrorString(_e)};         \
    } while (0)

}  // namespace

// ===========================================================================
// minimize_SQP (solver.cpp:109-197): contact as hard constraints.
//
// Every constraint row of the reference acts on one vertex: NeumannBC rows
// x_i - xPrev_i - dt v_i = 0 on masked DOFs (neumannBC.cpp:4-40), and
// PlaneContact / DiskContact rows n.(x + dx - p) >= 0 on active surface
// vertices (planeContact.cpp:4-84, diskContact.cpp:4-88).  The nodes that
// carry rows form the K table (sorted, one entry per node).  Each SQP
// iteration solves the QP  min 1/2 dx^T H dx + g^T dx  s.t. those rows by a
// primal-dual active set method over the inequality rows: for a working set
// W, the rows of each K node (its equality rows and its rows in W, linearly
// dependent ones dropped) give the node's particular step dxp (minimum norm),
// its projector P = I - A^T (A A^T)^-1 A and the projected PCG solve of
// the rest; then the multipliers (A A^T)^-1 A (H dx + g)_node decide which
// W rows leave (negative multiplier) and which violated rows join.  W starts
// from the rows violated at the step's start and carries over between SQP
// iterations; the reference solves the same QPs with OSQP.
// ===========================================================================
constexpr double kPlaneThreshold = 0.01;       // PLANE_CONSTRAINT_THRESHOLD (common.h:29)
constexpr double kDiskPlaneThreshold = 0.01;   // DISK_PLANAR_CONSTRAINT_THRESHOLD (common.h:33)
constexpr double kDiskRadialThreshold = 0.001; // DISK_RADIAL_CONSTRAINT_THRESHOLD (common.h:35)
constexpr int kSqpActiveThreads = 1024;

struct SqpDev {
    int nK, nsurf;
    const int* knode;    // [nK] node of entry k (ascending)
    const int* keq;      // [nK] masked DOFs of the node (bit c: DOF 3 node + c)
    const int* surf2k;   // [nsurf] entry of each surface vertex
    int* kpid;           // [nK] active plane of the node or -1
    int* kdid;           // [nK] captured disk of the node or -1
    int* kw;             // [nK] working set: bit 0 the plane row, bit 1 the disk row
    int* pmask;          // [nsurf] planeCollisionMask_
    int* dmask;          // [nsurf] diskCollisionMask_ (kept between steps)
    int* frozen;         // [1] frozenActiveSet_
    int* changes;        // [1] working-set changes of the last multiplier pass
    double* dxfull;      // [n] particular step (nonzero on K nodes only)
    double vtol;         // violation tolerance of a row outside W (m)
};

__device__ __forceinline__ double plane_dist_dev(const ContactArgs& c, int j, const double* x) {
    return (x[0] - c.pp[j][0]) * c.pn[j][0] + (x[1] - c.pp[j][1]) * c.pn[j][1] + (x[2] - c.pp[j][2]) * c.pn[j][2];
}

__device__ __forceinline__ double disk_plane_dist_dev(const ContactArgs& c, int d, const double* x) {
    return c.dn[d][0] * (x[0] - c.dc[d][0]) + c.dn[d][1] * (x[1] - c.dc[d][1]) + c.dn[d][2] * (x[2] - c.dc[d][2]);
}

__device__ __forceinline__ double disk_radial_dist_dev(const ContactArgs& c, int d, const double* x) {
    const double t = disk_plane_dist_dev(c, d, x);
    double r = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double w = (x[k] - c.dc[d][k]) - t * c.dn[d][k];
        r += w * w;
    }
    return sqrt(r);
}

// PlaneContact::update_active_constraints (planeContact.cpp:87-124) and
// DiskContact::update_active_constraints (diskContact.cpp:91-202) at the
// step's initial guess x, in one workgroup (the disk rule has two
// footprint-wide decisions); then the K table's active rows and the initial
// working set (rows violated at x).
__global__ __launch_bounds__(kSqpActiveThreads) void k_sqp_active(SqpDev s, ContactArgs c, const double* x) {
    __shared__ int sh_clear_fail, sh_hit, sh_cnt;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int i = tid; i < s.nsurf; i += nt) {
        const double* xv = x + 3 * (size_t)c.surf[i];
        double best = 0.0;
        int idx = -1;
        for (int j = 0; j < c.nplanes; ++j) {
            const double d = plane_dist_dev(c, j, xv);
            if (d < kPlaneThreshold && (idx < 0 || d < best)) {
                best = d;
                idx = j;
            }
        }
        s.pmask[i] = idx;
    }
    if (c.ndisks > 0) {
        const double pt = kDiskPlaneThreshold, rt = kDiskRadialThreshold;
        const double clear = 3.0 * pt, depth = 5.0 * pt;
        if (tid == 0) sh_clear_fail = sh_hit = sh_cnt = 0;
        __syncthreads();
        if (*s.frozen) {
            // frozen: keep the footprint until all of it is clearly above its disk
            for (int i = tid; i < s.nsurf; i += nt) {
                const int d = s.dmask[i];
                if (d >= 0 && disk_plane_dist_dev(c, d, x + 3 * (size_t)c.surf[i]) < clear)
                    atomicOr(&sh_clear_fail, 1);
            }
            __syncthreads();
            if (!sh_clear_fail) {
                for (int i = tid; i < s.nsurf; i += nt) s.dmask[i] = -1;
                if (tid == 0) *s.frozen = 0;
            }
        } else {
            for (int i = tid; i < s.nsurf; i += nt) {
                const double* xv = x + 3 * (size_t)c.surf[i];
                for (int d = 0; d < c.ndisks; ++d)
                    if (disk_plane_dist_dev(c, d, xv) < pt && !(disk_radial_dist_dev(c, d, xv) > c.dr[d] + rt))
                        atomicOr(&sh_hit, 1 << d);
            }
            __syncthreads();
            const int hit = sh_hit;
            for (int i = tid; i < s.nsurf; i += nt) {
                const double* xv = x + 3 * (size_t)c.surf[i];
                int slot = -1;
                double sabs = 0.0;
                for (int d = 0; d < c.ndisks; ++d) {
                    if (!((hit >> d) & 1)) continue;
                    const double pd = disk_plane_dist_dev(c, d, xv);
                    if (pd > pt || pd < -depth) continue;
                    if (disk_radial_dist_dev(c, d, xv) > c.dr[d] + rt) continue;
                    if (slot < 0 || fabs(pd) < sabs) {
                        slot = d;
                        sabs = fabs(pd);
                    }
                }
                s.dmask[i] = slot;
                if (slot >= 0) atomicAdd(&sh_cnt, 1);
            }
            __syncthreads();
            if (tid == 0 && sh_cnt > 0) *s.frozen = 1;
        }
    }
    __syncthreads();
    for (int i = tid; i < s.nsurf; i += nt) {
        const int k = s.surf2k[i];
        const double* xv = x + 3 * (size_t)c.surf[i];
        const int pid = s.pmask[i], did = c.ndisks > 0 ? s.dmask[i] : -1;
        int w = 0;
        if (pid >= 0 && -plane_dist_dev(c, pid, xv) > 0.0) w |= 1;
        if (did >= 0 && -disk_plane_dist_dev(c, did, xv) > 0.0) w |= 2;
        s.kpid[k] = pid;
        s.kdid[k] = did;
        s.kw[k] = w;
    }
}

// The rows of K entry k for working set w: equality rows first (unit
// vectors of the masked DOFs), then the plane row and the disk row when in w;
// a row linearly dependent on the rows before it is dropped (its bit cleared
// in *w).  Builds the Cholesky factor L of the Gram matrix A A^T of the kept
// rows.  b: each row's right-hand side (a.dx >= b or = b).
struct SqpRows {
    int nr;          // kept rows
    int kind[3];     // 0 equality, 1 plane, 2 disk
    double a[3][3];
    double b[3];
    double L[3][3];  // lower triangular, G = L L^T
};

__device__ bool sqp_try_row(SqpRows& R, const double* a, double b, int kind) {
    if (R.nr == 3) return false;
    double y[3];
    const double aa = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    double d = aa;
    for (int j = 0; j < R.nr; ++j) {
        double gj = R.a[j][0] * a[0] + R.a[j][1] * a[1] + R.a[j][2] * a[2];
        for (int i = 0; i < j; ++i) gj -= R.L[j][i] * y[i];
        y[j] = gj / R.L[j][j];
        d -= y[j] * y[j];
    }
    if (!(d > 1e-10 * aa)) return false;
    const int r = R.nr;
    for (int j = 0; j < r; ++j) R.L[r][j] = y[j];
    R.L[r][r] = sqrt(d);
    for (int k = 0; k < 3; ++k) R.a[r][k] = a[k];
    R.b[r] = b;
    R.kind[r] = kind;
    R.nr = r + 1;
    return true;
}

__device__ void sqp_rows(const SqpDev& s, const ContactArgs& c, int k, int* w, const double* x, const double* xprev,
                         const double* bcval, double dt, SqpRows& R) {
    R.nr = 0;
    const int node = s.knode[k], eq = s.keq[k];
    const double* xv = x + 3 * (size_t)node;
    for (int comp = 0; comp < 3; ++comp) {
        if (!((eq >> comp) & 1)) continue;
        const size_t i = 3 * (size_t)node + comp;
        const double a[3] = {comp == 0 ? 1.0 : 0.0, comp == 1 ? 1.0 : 0.0, comp == 2 ? 1.0 : 0.0};
        sqp_try_row(R, a, -(x[i] - xprev[i] - dt * bcval[i]), 0);
    }
    const int pid = s.kpid[k], did = s.kdid[k];
    if (pid >= 0 && (*w & 1) && !sqp_try_row(R, c.pn[pid], -plane_dist_dev(c, pid, xv), 1)) *w &= ~1;
    if (did >= 0 && (*w & 2) && !sqp_try_row(R, c.dn[did], -disk_plane_dist_dev(c, did, xv), 2)) *w &= ~2;
}

// G^-1 t for the kept rows (two triangular solves with L)
__device__ void sqp_gram_solve(const SqpRows& R, const double* t, double* out) {
    double y[3];
    for (int j = 0; j < R.nr; ++j) {
        double v = t[j];
        for (int i = 0; i < j; ++i) v -= R.L[j][i] * y[i];
        y[j] = v / R.L[j][j];
    }
    for (int j = R.nr - 1; j >= 0; --j) {
        double v = y[j];
        for (int i = j + 1; i < R.nr; ++i) v -= R.L[i][j] * out[i];
        out[j] = v / R.L[j][j];
    }
}

// block-Jacobi rows of the K nodes, kept before their projection
template <typename HT>
__global__ __launch_bounds__(kBlock) void k_sqp_save_minv(SqpDev s, const HT* Minv, HT* Msave) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= s.nK) return;
    const HT* M = Minv + 9 * (size_t)s.knode[k];
#pragma unroll
    for (int i = 0; i < 9; ++i) Msave[9 * (size_t)k + i] = M[i];
}

// per K entry for the working set: dxp = A^T G^-1 b (into dxfull), P = I -
// A^T G^-1 A, and the node's block-Jacobi rows P D^-1 P
template <typename HT>
__global__ __launch_bounds__(kBlock) void k_sqp_setup(SqpDev s, ContactArgs c, const double* x, const double* xprev,
                                                      const double* bcval, double dt, const HT* Msave, HT* P,
                                                      HT* Minv) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= s.nK) return;
    int w = s.kw[k];
    SqpRows R;
    sqp_rows(s, c, k, &w, x, xprev, bcval, dt, R);
    s.kw[k] = w;
    double yb[3] = {0.0, 0.0, 0.0};
    sqp_gram_solve(R, R.b, yb);
    // A^T G^-1 A column by column: G^-1 (A e_m) for m = 0..2
    double Pm[9];
    for (int m = 0; m < 3; ++m) {
        double t[3], z[3] = {0.0, 0.0, 0.0};
        for (int j = 0; j < R.nr; ++j) t[j] = R.a[j][m];
        sqp_gram_solve(R, t, z);
        for (int l = 0; l < 3; ++l) {
            double v = 0.0;
            for (int j = 0; j < R.nr; ++j) v += R.a[j][l] * z[j];
            Pm[3 * l + m] = (l == m ? 1.0 : 0.0) - v;
        }
    }
    const int node = s.knode[k];
    for (int l = 0; l < 3; ++l) {
        double v = 0.0;
        for (int j = 0; j < R.nr; ++j) v += R.a[j][l] * yb[j];
        s.dxfull[3 * (size_t)node + l] = v;
    }
    const HT* Ms = Msave + 9 * (size_t)k;
    double MP[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double v = 0.0;
            for (int l = 0; l < 3; ++l) v += (double)Ms[3 * i + l] * Pm[3 * l + j];
            MP[3 * i + j] = v;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double v = 0.0;
            for (int l = 0; l < 3; ++l) v += Pm[3 * i + l] * MP[3 * l + j];
            Minv[9 * (size_t)node + 3 * i + j] = (HT)v;
            P[9 * (size_t)k + 3 * i + j] = (HT)Pm[3 * i + j];
        }
}

// (H x)_node on the sliced node-block layout, FP64 accumulation; x = xa + xb
template <typename HT, typename XB>
__device__ __forceinline__ void blk_row_apply(const int* sbase, const int* bcol, const HT* vals, int node,
                                              const double* xa, const XB* xb, double (&out)[3]) {
    const int sl = node / kSlice, lane = node % kSlice;
    const int base = sbase[sl];
    const int width = (sbase[sl + 1] - base) / kSlice;
    out[0] = out[1] = out[2] = 0.0;
    for (int j = 0; j < width; ++j) {
        const int nbr = bcol[base + kSlice * j + lane];
        double xx[3];
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) xx[cc] = xa[3 * (size_t)nbr + cc] + (xb ? (double)xb[3 * (size_t)nbr + cc] : 0.0);
        const HT* bb = vals + 9LL * base + lane + (long long)kSlice * 9 * j;
#pragma unroll
        for (int r = 0; r < 3; ++r)
            out[r] += (double)bb[(3 * r) * kSlice] * xx[0] + (double)bb[(3 * r + 1) * kSlice] * xx[1] +
                      (double)bb[(3 * r + 2) * kSlice] * xx[2];
    }
}

// rhs = -(g + H dxp): the QP's residual at its particular step
template <typename HT>
__global__ __launch_bounds__(kBlock) void k_sqp_rhs(int N, const int* sbase, const int* bcol, const HT* vals,
                                                    const double* dxp, const double* g, double* rhs) {
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= N) return;
    double hx[3];
    blk_row_apply<HT, double>(sbase, bcol, vals, node, dxp, (const double*)nullptr, hx);
#pragma unroll
    for (int r = 0; r < 3; ++r) rhs[3 * (size_t)node + r] = -(g[3 * (size_t)node + r] + hx[r]);
}

// Multipliers and the next working set.  dx = dxp + y (y: the projected
// solve, HT); per K node s = (H dx + g)_node and the kept rows' multipliers
// G^-1 A s.  A working-set row with a negative multiplier (below -1e-6
// |s|_inf) leaves; an active row outside it that dx violates by more than
// vtol, and is independent of the node's kept rows, joins.  Counts the
// changes into *changes.
template <typename HT>
__global__ __launch_bounds__(kBlock) void k_sqp_multipliers(SqpDev s, ContactArgs c, const int* sbase,
                                                            const int* bcol, const HT* vals, const HT* y,
                                                            const double* g, const double* x, const double* xprev,
                                                            const double* bcval, double dt) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= s.nK) return;
    const int node = s.knode[k];
    int w = s.kw[k];
    SqpRows R;
    sqp_rows(s, c, k, &w, x, xprev, bcval, dt, R);
    double hx[3];
    blk_row_apply<HT, HT>(sbase, bcol, vals, node, s.dxfull, y, hx);
    double sv[3], dxv[3];
    for (int r = 0; r < 3; ++r) {
        sv[r] = g[3 * (size_t)node + r] + hx[r];
        dxv[r] = s.dxfull[3 * (size_t)node + r] + (double)y[3 * (size_t)node + r];
    }
    const double smax = fmax(fabs(sv[0]), fmax(fabs(sv[1]), fabs(sv[2])));
    double t[3], lam[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < R.nr; ++j) t[j] = R.a[j][0] * sv[0] + R.a[j][1] * sv[1] + R.a[j][2] * sv[2];
    sqp_gram_solve(R, t, lam);
    int nw = w, chg = 0;
    for (int j = 0; j < R.nr; ++j) {
        if (R.kind[j] == 0 || !(lam[j] < -1e-6 * smax)) continue;
        nw &= ~(R.kind[j] == 1 ? 1 : 2);
        ++chg;
    }
    // rows outside the working set: join when violated and independent
    const double* xv = x + 3 * (size_t)node;
    const int pid = s.kpid[k], did = s.kdid[k];
    for (int kind = 1; kind <= 2; ++kind) {
        const int bit = kind == 1 ? 1 : 2;
        const int id = kind == 1 ? pid : did;
        if (id < 0 || (w & bit)) continue;
        const double* a = kind == 1 ? c.pn[id] : c.dn[id];
        const double b = kind == 1 ? -plane_dist_dev(c, id, xv) : -disk_plane_dist_dev(c, id, xv);
        if (!(a[0] * dxv[0] + a[1] * dxv[1] + a[2] * dxv[2] < b - s.vtol)) continue;
        if (!sqp_try_row(R, a, b, kind)) continue;
        nw |= bit;
        ++chg;
    }
    s.kw[k] = nw;
    if (chg) atomicAdd(s.changes, chg);
}

// dx = dxp + y (FP64)
template <typename HT>
__global__ void k_sqp_step(long long n, const double* dxp, const HT* y, double* dx) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dx[i] = dxp[i] + (double)y[i];
}

__global__ __launch_bounds__(kBlock) void k_sumsq(int n, const double* v, Reduce red) {
    double acc[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) acc[0] += v[i] * v[i];
    block_partial<1>(acc, red.partials);
}

struct sors_ctx {
    int N = 0, E = 0, ed = 0, n = 0;
    int material = SORS_NEOHOOKEAN, scheme = SORS_BACKWARD_EULER, method = SORS_MINIMIZE_NEWTON;
    double grav[3] = {0, 0, 0};
    double damping = 0;
    bool has_mask = false, has_fext = false;
    sors_options opt;
    std::string err;
    hipStream_t stream = nullptr;
    // host copies needed for patterns / exports
    std::vector<int> h_conn;   // [E][ed]
    std::vector<int> h_mask;   // [3N]
    std::vector<double> h_rest;  // [3N] rest positions (coarse-space aggregation)
    // device buffers
    DevMem conn, rest, mu, lam, mass, mask, bcval;
    DevMem q, qprev, vprev, gprev, gtmp, gint, g, fext, dxd, qtmp;
    DevMem rowptr, cols, diag, rank, fullmap, vals, Minv;
    DevMem pr, pu, pp, ps, pw, px;
    DevMem partials, scal, pcgst, pcgctl, pcgpart, flag;
    // a serving loop points the result scalars, the element flag and the PCG
    // outcome record of its rollouts into one slot array it reads at once
    double* scal_ext = nullptr;
    int* flag_ext = nullptr;
    double* S() const { return scal_ext ? scal_ext : scal.as<double>(); }
    int* F() const { return flag_ext ? flag_ext : flag.as<int>(); }
    double* h_pin = nullptr;   // pinned readback area
    HostPattern pat;
    int pat_dtnz = -1;
    bool pat_on_device = false;
    DevMem pcgst_bench, copybuf;
    DevMem sbase, bcol, selfslot;   // sliced node-block layout (node-granular masks)
    DevMem bcol16;                  // bcol as 16-bit offsets from the row's node (when they all fit)
    bool col16 = false;
    DevMem incptr, inc;             // node -> (element, local node) incidence (gradient sums)
    DevMem teg, teb;                // element rows between the two assembly phases
    DevMem tsptr, tsidx, tqslice;   // layout slot -> element blocks (pattern.h)
    DevMem vold;                    // velocity before the last step (second-order predictor)
    // two-level preconditioner (options.coarse_nodes > 0)
    int coarse_built_for = 0;       // coarse_nodes of the built space (0: none)
    int coarse_gen = 0;             // bumped on every rebuild (captured graphs hold its buffers)
    long long coarse_tangents = 0;  // tangents seen since the coarse space was built
    hipStream_t side = nullptr;     // coarse inverse (default priority, like the main stream)
    hipEvent_t evG = nullptr, evR = nullptr;  // E assembled / inversion done
    hipEvent_t evPre = nullptr;     // served rollouts: preconditioner built (the batched PCG waits)
    bool refresh_pending = false;   // an inversion on the side stream not yet collected
    const void* coarse_pat = nullptr;
    bool coarse_ready = false;      // E^-1 valid for the current matrix
    int nagg = 0, nc = 0, npairs = 0, ngchunks = 0;
    int npad = 0;                   // nc rounded up to the Gauss-Jordan tile
    DevMem cdn, captr, canodes, crot, cpairI, cpairJ, cpchunk, cgptr, cci, ccj, ccbase;
    DevMem cE, cEinv, cX, cY, crc, cgpart, cpart, cinfo;   // cX / cY: Gauss-Jordan ping-pong (npad^2)
    DevMem cEinv32;                 // FP32 copy of E^-1 (the mixed-precision apply)
    DevMem cpinv;                   // pivot block inverses of the current and the next step (2 x 32 x 32)
    const double* cres = nullptr;   // the buffer holding E^-1 after the last elimination
    bool vold_valid = false;
    double vold_dt = 0.0;
    // active fibre stress (sors_set_active) and parameter adjoints
    bool has_active = false;
    int ngroups = 0;
    double sigma = 0.0;
    std::vector<double> h_nu;       // Poisson ratio per element (Lame updates, dE adjoint)
    DevMem agroup, afib, act, nu, adjb, adjq, adjv, adjE, adjA;
    // SoftCon muscle (sors_set_muscle)
    bool has_muscle = false;
    int mngroups = 0;
    DevMem mgroup, mdir, mk, mact;
    // PressureForce (sors_set_pressure): faces, groups, pressures, the
    // node -> face incidence, face forces and the node forces at q (fp0)
    // and at the line-search point (fp1)
    bool has_pressure = false;
    bool export_pressure_jacobian = false;  // set by sors_hessian alone
    int pnf = 0, pnv = 0, pngroups = 0;
    DevMem pfaces, pgroup, pact, pincptr, pinc, pff, fp0, fp1;
    // contact (sors_set_contact): primitives by value, candidates and lag on the device
    bool has_contact = false;
    ContactArgs cargs{};
    DevMem csurf, lagN, lagn, lagD, gcon;
    long long nvals = 0;            // value slots of the active layout
    // captured PCG iteration chunk (one per precision)
    // captured PCG chunks per precision and chunk size (pcg_check_every / 4, / 2, / 1)
    hipGraphExec_t gexec[2][3] = {};
    const void* gkey[2][3] = {};
    int gchunk[2][3] = {};
    int gcoarse[2][3] = {};
    bool block() const { return pat.node_granular; }
    // per-launch kernel timing (options.profile_kernels)
    sors_stats* prof = nullptr;
    std::vector<hipEvent_t> pev;
    hipEvent_t* events(int count) {
        while ((int)pev.size() < count) {
            hipEvent_t e;
            // timing events of kernels whose results the next kernel reads: a
            // device-scope release (the per-XCD L2s are written back, as after
            // any launch) and no system-scope one; the disable-system-fence
            // flag skips the cache writeback altogether
            HIPX(hipEventCreateWithFlags(&e, hipEventReleaseToDevice));
            pev.push_back(e);
        }
        return pev.data();
    }

    Reduce red() { return Reduce{partials.as<double>()}; }
    PcgPart pcg_part() {
        double* b = pcgpart.as<double>();
        return PcgPart{{b, b + 2 * kMaxRedBlocks}, b + 4 * kMaxRedBlocks};
    }
    bool sqp() const { return method == SORS_MINIMIZE_SQP; }
    bool pat_sqp = false;           // the device pattern has the SQP (unmasked) semantics
    // minimize_SQP: the K table of constrained nodes (Dirichlet DOFs, contact
    // candidates), their active and working sets, projectors and steps
    bool sqp_built = false;
    int nK = 0;
    double sqp_vtol = 0.0;
    std::vector<int> h_surf;        // contact candidates (host copy)
    DevMem sknode, skeq, ssurf2k, skpid, skdid, skw, skmap, spmask, sdmask, sfrozen, schg;
    DevMem sP, sMsave, sdxfull, srhs;
    // this device's static scratch (sors_energy_noalloc), set up by sors_create
    double* scratch_dev = nullptr;
    std::string scratch_err;
};

namespace {

// per-element material constants held on the device: the Lame constants of
// (E, nu) (neoHookean.cpp:10-11), reparameterised for Stable Neo-Hookean as
// mu' = 4/3 mu, lam' = lam + 5/8 mu' (stableNeoHookean.cpp:9-14)
void material_constants(int material, double Ey, double nu, double* mu, double* lam) {
    *lam = (Ey * nu) / ((1 + nu) * (1 - 2 * nu));
    *mu = Ey / (2 * (1 + nu));
    if (material == SORS_STABLE_NEOHOOKEAN) {
        *mu = (4.0 / 3.0) * (*mu);
        *lam = *lam + (5.0 / 8.0) * (*mu);
    }
}

int fail(sors_ctx* c, int code, const std::string& m) {
    if (c) c->err = m;
    return code;
}

template <typename T>
void upload(DevMem& d, const T* h, size_t count, hipStream_t st) {
    if (!d.alloc(sizeof(T) * count)) throw HipError{"hipMalloc failed (" + std::to_string(sizeof(T) * count) + " bytes)"};
    if (count) HIPX(hipMemcpyAsync(d.p, h, sizeof(T) * count, hipMemcpyHostToDevice, st));
}

// a host vector, which may be empty: at least one element is allocated so
// that the device pointer is valid, and only the vector's own elements are copied
template <typename T>
void upload_vec(DevMem& d, const std::vector<T>& v, hipStream_t st) {
    if (!d.alloc(sizeof(T) * std::max<size_t>(v.size(), 1)))
        throw HipError{"hipMalloc failed (" + std::to_string(sizeof(T) * v.size()) + " bytes)"};
    if (!v.empty()) HIPX(hipMemcpyAsync(d.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, st));
}

void ensure(DevMem& d, size_t bytes) {
    if (!d.alloc(bytes)) throw HipError{"hipMalloc failed (" + std::to_string(bytes) + " bytes)"};
}

int grid_for(long long work, int block = kBlock) {
    long long b = (work + block - 1) / block;
    return (int)std::max<long long>(1, std::min<long long>(b, 1 << 30));
}

void sync(sors_ctx* c) { HIPX(hipStreamSynchronize(c->stream)); }

// The SpMV's column table as 16-bit offsets from each row's node, when every
// offset fits (node orderings with a bandwidth under 32,768 nodes, such as
// the structured beams' 1,723); padding lanes past the last node point at
// the last node (their blocks are zero and their rows are not stored)
#ifndef SORS_COL16
#define SORS_COL16 1
#endif
void build_col16(sors_ctx* c) {
    c->col16 = false;
    if (!SORS_COL16) return;
    const int ns = c->pat.nslices;
    const long long slots = c->pat.sbase[ns];
    if (c->N == 0 || slots == 0) return;
    std::vector<short> off((size_t)slots);
    for (int sl = 0; sl < ns; ++sl) {
        const int base = c->pat.sbase[sl], w = (c->pat.sbase[sl + 1] - base) / kSlice;
        for (int lane = 0; lane < kSlice; ++lane) {
            const int v = sl * kSlice + lane;
            for (int k = 0; k < w; ++k) {
                const long long p = base + (long long)kSlice * k + lane;
                const int d = v < c->N ? c->pat.bcol[(size_t)p] - v : (c->N - 1) - v;
                if (d < -32768 || d > 32767) return;
                off[(size_t)p] = (short)d;
            }
        }
    }
    upload_vec(c->bcol16, off, c->stream);
    c->col16 = true;
}

// (re)build the pattern for dt == 0 / dt != 0 and upload it
void ensure_pattern(sors_ctx* c, double dt) {
    const int dtnz = dt != 0.0;
    if (c->pat_dtnz == dtnz && c->pat_sqp == c->sqp() && c->pat_on_device) return;
    std::string msg;
    if (!build_pattern(c->N, c->E, c->ed, c->h_conn.data(), c->has_mask ? c->h_mask.data() : nullptr, dtnz, c->sqp(),
                       c->pat, &msg))
        throw HipError{"pattern: " + msg};
    upload(c->rowptr, c->pat.rowptr.data(), c->pat.rowptr.size(), c->stream);
    upload(c->cols, c->pat.cols.data(), c->pat.cols.size(), c->stream);
    if (c->pat.node_granular) {
        upload(c->rank, c->pat.rank.data(), c->pat.rank.size(), c->stream);
        upload(c->sbase, c->pat.sbase.data(), c->pat.sbase.size(), c->stream);
        upload(c->bcol, c->pat.bcol.data(), c->pat.bcol.size(), c->stream);
        build_col16(c);
        upload(c->selfslot, c->pat.self_slot.data(), c->pat.self_slot.size(), c->stream);
        upload(c->diag, c->pat.bdiag.data(), c->pat.bdiag.size(), c->stream);
        c->nvals = 9LL * c->pat.sbase[c->pat.nslices];
        if (!c->pat.sptr.empty()) {
            upload(c->incptr, c->pat.incptr.data(), c->pat.incptr.size(), c->stream);
            upload(c->inc, c->pat.inc.data(), c->pat.inc.size(), c->stream);
            upload(c->tsptr, c->pat.sptr.data(), c->pat.sptr.size(), c->stream);
            upload_vec(c->tsidx, c->pat.sidx, c->stream);
            upload_vec(c->tqslice, c->pat.qslice, c->stream);
        }
    } else {
        upload(c->fullmap, c->pat.fullmap.data(), c->pat.fullmap.size(), c->stream);
        upload(c->diag, c->pat.diag.data(), c->pat.diag.size(), c->stream);
        c->nvals = c->pat.nnz;
    }
    ensure(c->vals, sizeof(double) * std::max<long long>(c->nvals, 1));
    for (auto& row : c->gkey)
        for (auto& k : row) k = nullptr;  // captured graphs hold the old buffers
    sync(c);  // host vectors of the pattern stay alive in c->pat
    c->pat_dtnz = dtnz;
    c->pat_sqp = c->sqp();
    c->pat_on_device = true;
}

ElemArgs elem_args(sors_ctx* c, const double* q) {
    ElemArgs a{};
    a.E = c->E;
    a.ed = c->ed;
    a.N = c->N;
    a.conn = c->conn.as<int>();
    a.rest = c->rest.as<double>();
    a.q = q;
    a.dq = nullptr;
    a.alpha = 0.0;
    a.mu = c->mu.as<double>();
    a.lam = c->lam.as<double>();
    a.mask = c->has_mask ? c->mask.as<int>() : nullptr;
    a.sqp = c->sqp();
    a.material = c->material;
    a.hscale = 1.0;
    a.rowptr = c->rowptr.as<int>();
    a.rank = c->pat.node_granular ? c->rank.as<uint8_t>() : nullptr;
    a.fullmap = c->pat.node_granular ? nullptr : c->fullmap.as<int>();
    a.sbase = c->block() ? c->sbase.as<int>() : nullptr;
    a.flag = c->F();
    if (c->has_active) {
        a.agroup = c->agroup.as<int>();
        a.afib = c->afib.as<double>();
        a.act = c->act.as<double>();
        a.sigma = c->sigma;
    }
    if (c->has_muscle) {
        a.mgroup = c->mgroup.as<int>();
        a.mdir = c->mdir.as<double>();
        a.mk = c->mk.as<double>();
        a.mact = c->mact.as<double>();
    }
    return a;
}

PressArgs press_args(sors_ctx* c) {
    PressArgs a;
    a.nf = c->pnf;
    a.nv = c->pnv;
    a.N = c->N;
    a.faces = c->pfaces.as<int>();
    a.group = c->pgroup.as<int>();
    a.pact = c->pact.as<double>();
    a.incptr = c->pincptr.as<int>();
    a.inc = c->pinc.as<int>();
    return a;
}

// node pressure forces at q + alpha dq (dq may be null) into fp
void pressure_force(sors_ctx* c, const double* q, const double* dq, double alpha, double* fp, hipStream_t st) {
    const PressArgs a = press_args(c);
    k_pressure_faces<<<grid_for(a.nf, kBlock), kBlock, 0, st>>>(a, q, dq, alpha, c->pff.as<double>());
    k_pressure_nodes<<<grid_for(a.N, kBlock), kBlock, 0, st>>>(a, c->pff.as<double>(), fp);
}

// contact enters the objective as the penalty terms (minimize_newton[_linesearch],
// frozen Hessian); under minimize_SQP it is a set of constraints
bool penalty_on(const sors_ctx* c) { return c->has_contact && c->cargs.nsurf > 0 && !c->sqp(); }

NodeArgs node_args(sors_ctx* c, const double* q, double dt) {
    NodeArgs a{};
    if (c->has_pressure) a.fpres = c->fp0.as<double>();
    a.n = c->n;
    a.dt = dt;
    a.scheme = c->scheme;
    a.sqp = c->sqp();
    a.damping = c->damping;
    for (int k = 0; k < 3; ++k) a.grav[k] = c->grav[k];
    a.mass = c->mass.as<double>();
    a.q = q;
    a.qprev = c->qprev.as<double>();
    a.vprev = c->vprev.as<double>();
    a.gprev = c->gprev.as<double>();
    a.gint = c->gint.as<double>();
    a.fext = c->has_fext ? c->fext.as<double>() : nullptr;
    a.mask = c->has_mask ? c->mask.as<int>() : nullptr;
    a.bcval = c->bcval.as<double>();
    a.g = c->g.as<double>();
    a.gtmp = c->gtmp.as<double>();
    a.diag = c->diag.as<int>();
    a.gcon = penalty_on(c) ? c->gcon.as<double>() : nullptr;
    return a;
}

ContactArgs contact_args(sors_ctx* c) {
    ContactArgs a = c->cargs;
    a.surf = c->csurf.as<int>();
    a.mask = c->has_mask ? c->mask.as<int>() : nullptr;
    a.lagN = c->lagN.as<double>();
    a.lagn = c->lagn.as<double>();
    a.lagD = c->lagD.as<double>();
    return a;
}

void contact_lag(sors_ctx* c) {
    if (!c->has_contact || c->cargs.nsurf == 0) return;
    k_contact_lag<<<grid_for(c->cargs.nsurf), kBlock, 0, c->stream>>>(contact_args(c), c->qprev.as<double>());
    HIPX(hipGetLastError());
}

// tet4 / hex8 Neo-Hookean or Stable Neo-Hookean on the sliced layout: the
// two-phase assembly writes every gradient entry and every value of the
// layout itself (no fills, no atomics)
bool node_assembly(const sors_ctx* c) {
    return c->block() && (c->ed == 4 || c->ed == 8) &&
           (c->material == SORS_NEOHOOKEAN || c->material == SORS_STABLE_NEOHOOKEAN) && !c->pat.sptr.empty();
}

// element rows between the assembly phases (hbytes: bytes per Hessian value, 0: gradient only)
void ensure_elem_rows(sors_ctx* c, size_t hbytes) {
    if (!node_assembly(c)) return;
    const size_t Ep = (size_t)elem_rows_padded(c->E);  // rows are tiled by 64 elements
    ensure(c->teg, sizeof(double) * 4 * c->ed * Ep);
    if (hbytes)
        ensure(c->teb, hbytes * (hbytes == 4 ? tet_blk<float>() : tet_blk<double>()) * elem_blocks(c->ed) * Ep);
}

// element kernel launch; with e0/e1 the dispatches record the two events
template <typename HT, int MODE>
void launch_elem(sors_ctx* c, const ElemArgs& a, int project, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    if (MODE != 2 && node_assembly(c)) {
        if (c->E == 0) {
            fill<double>(c->gint.as<double>(), c->n, 0.0, c->stream);
            if (MODE == 0) fill<HT>((HT*)c->vals.p, c->nvals, (HT)0, c->stream);
            return;
        }
        // element rows (phase 1), then the per-slot / per-node sums into the layout (phase 2)
        ensure_elem_rows(c, MODE == 0 ? sizeof(HT) : 0);
        double* eg = c->teg.as<double>();
        HT* eb = MODE == 0 ? (HT*)c->teb.p : nullptr;
        // Stable Neo-Hookean or a muscle: the general element kernels
        const bool gen = c->material == SORS_STABLE_NEOHOOKEAN || c->has_muscle;
        if (c->ed == 4) {
            void (*k1)(ElemArgs, double*, HT*) =
                gen ? (project ? k_tet_elem<HT, 1, MODE, 1> : k_tet_elem<HT, 0, MODE, 1>)
                    : (project ? k_tet_elem<HT, 1, MODE, 0> : k_tet_elem<HT, 0, MODE, 0>);
            hipExtLaunchKernelGGL(k1, dim3(grid_for(c->E, kBlock)), dim3(kBlock), 0, c->stream, e0, nullptr, 0, a,
                                  eg, eb);
        } else {
            hipExtLaunchKernelGGL(gen ? k_hex_grad<1> : k_hex_grad<0>, dim3(grid_for(c->E, 128)), dim3(128), 0,
                                  c->stream, e0, nullptr, 0, a, eg);
            if (MODE == 0) {
                void (*kr)(ElemArgs, HT*) = gen ? (project ? k_hex_rows<HT, 1, 1> : k_hex_rows<HT, 0, 1>)
                                                : (project ? k_hex_rows<HT, 1, 0> : k_hex_rows<HT, 0, 0>);
                kr<<<grid_for(8LL * c->E, kHexRowThreads), kHexRowThreads, 0, c->stream>>>(a, eb);
            }
        }
        const int* incptr = c->incptr.as<int>();
        const int* inc = c->inc.as<int>();
        if (MODE == 0) {
            const int nq = c->pat.sbase[c->pat.nslices] / kSlice;
            auto k2 = c->ed == 4 ? k_slot_sums<HT, true> : k_slot_sums<HT, false>;
            hipExtLaunchKernelGGL(k2, dim3(grid_for((long long)nq * kSlice, kBlock)), dim3(kBlock), 0,
                                  c->stream, nullptr, e1, 0, a, nq, (const int*)c->tqslice.as<int>(),
                                  (const int*)c->tsptr.as<int>(), (const int*)c->tsidx.as<int>(), (const HT*)eb,
                                  (const int*)c->selfslot.as<int>(), incptr, inc, (const double*)eg);
        } else {
            hipExtLaunchKernelGGL(k_grad_gather, dim3(grid_for(c->N, kBlock)), dim3(kBlock), 0, c->stream, nullptr,
                                  e1, 0, a, incptr, inc, (const double*)eg);
        }
        HIPX(hipGetLastError());
        return;
    }
    if (c->E == 0) return;
    const bool tet = c->ed == 4;
    const int block = tet ? kBlock : 128;
    const dim3 grid(grid_for(c->E, block));
    void (*k)(ElemArgs);
    if (tet)
        k = project ? k_tet<HT, 1, MODE> : k_tet<HT, 0, MODE>;
    else
        k = project ? k_hex<HT, 1, MODE> : k_hex<HT, 0, MODE>;
    if (e0)
        hipExtLaunchKernelGGL(k, grid, dim3(block), 0, c->stream, e0, e1, 0, a);
    else
        k<<<grid, block, 0, c->stream>>>(a);
    HIPX(hipGetLastError());
}

int read_flag(sors_ctx* c) {
    int f = 0;
    HIPX(hipMemcpyAsync(&f, c->flag.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    sync(c);
    return f;
}

void clear_flag(sors_ctx* c) { fill<int>(c->F(), 1, 0, c->stream); }

// gradient (FP64) and, when HT != void, the assembled Hessian; returns ||g||^2 in scal[S_GG]
template <typename HT>
void assemble(sors_ctx* c, const double* q, double dt, bool hessian, int project) {
    ElemArgs a = elem_args(c, q);
    if (!node_assembly(c)) {
        fill<double>(c->gint.as<double>(), c->n, 0.0, c->stream);
        if (hessian) fill<HT>((HT*)c->vals.p, c->nvals, (HT)0, c->stream);
    }
    a.gint = c->gint.as<double>();
    a.vals = hessian ? c->vals.p : nullptr;
    a.hscale = (c->scheme == SORS_CRANK_NICOLSON && dt != 0.0) ? 0.5 : 1.0;
    hipEvent_t* ev = c->prof && hessian ? c->events(2) : nullptr;
    if (hessian)
        launch_elem<HT, 0>(c, a, project, ev ? ev[0] : nullptr, ev ? ev[1] : nullptr);
    else
        launch_elem<HT, 1>(c, a, 0, ev ? ev[0] : nullptr, ev ? ev[1] : nullptr);
    if (ev) {
        HIPX(hipEventSynchronize(ev[1]));
        float ms = 0;
        HIPX(hipEventElapsedTime(&ms, ev[0], ev[1]));
        c->prof->assemble_kernel_ms += ms;
        c->prof->assemble_launches += 1;
    }
    if (c->has_pressure) {
        pressure_force(c, q, nullptr, 0.0, c->fp0.as<double>(), c->stream);
        // the force Jacobian (minimize_newton, energy.cpp:354-372) is part of
        // the exported Hessian; the solve's SPD tangent leaves it out (it is
        // not symmetric), so Newton's iterates differ but its fixed point does not
        if (hessian && c->export_pressure_jacobian && c->method == SORS_MINIMIZE_NEWTON) {
            const int* mk = c->has_mask ? c->mask.as<int>() : nullptr;
            k_pressure_jacobian<HT><<<grid_for(c->N, kBlock), kBlock, 0, c->stream>>>(
                press_args(c), q, c->sbase.as<int>(), c->bcol.as<int>(), c->selfslot.as<int>(), mk, a.hscale,
                (HT*)c->vals.p, c->F());
        }
    }
    if (penalty_on(c)) {
        const bool blk = c->block();
        k_contact_terms<HT><<<grid_for(c->cargs.nsurf), kBlock, 0, c->stream>>>(
            contact_args(c), q, c->qprev.as<double>(), dt, c->gcon.as<double>(), hessian ? (HT*)c->vals.p : nullptr,
            blk ? c->sbase.as<int>() : nullptr, blk ? c->selfslot.as<int>() : nullptr, c->rowptr.as<int>(),
            c->cols.as<int>());
    }
    NodeArgs na = node_args(c, q, dt);
    na.vals = hessian ? c->vals.p : nullptr;
    const int grid = red_grid(c->n, kBlock);
    k_node_terms<HT><<<grid, kBlock, 0, c->stream>>>(na, c->S() + S_GG, c->red());
    k_finalize<1><<<1, kBlock, 0, c->stream>>>(c->partials.as<double>(), grid, c->S() + S_GG);
    HIPX(hipGetLastError());
}

// ---- two-level preconditioner: set-up ---------------------------------------
CoarseArgs coarse_args(sors_ctx* c) {
    CoarseArgs a;
    a.nagg = c->nagg;
    a.nc = c->nc;
    a.dn = (const float4*)c->cdn.p;
    a.aptr = c->captr.as<int>();
    a.anodes = c->canodes.as<int>();
    a.rc = c->crc.as<double>();
    return a;
}

// nodes per aggregate: options.coarse_nodes, or fewer on a mesh of under
// 128 x coarse_nodes nodes, which keeps about 128 aggregates (64 nodes each
// at least; a 2,000-node mesh with 2,048-node aggregates would get one
// aggregate, whose six rigid modes barely help).  On the seeded serving
// stream (2k-200k nodes) 128 aggregates served 13.3 solves/s against 12.9
// with 32, 13.1 with 64 and 12.3 with 256; the 384 k-tet contact poke ran
// 11.3 against 11.9 ms per step with 32
int coarse_target(int N, int coarse_nodes) { return std::min(coarse_nodes, std::max(64, N / 128)); }

// (re)build the coarse space for the current pattern and options
void ensure_coarse(sors_ctx* c) {
    // minimize_SQP's projected solves run on block-Jacobi alone
    const int want = (c->opt.coarse_nodes > 0 && c->block() && !c->sqp()) ? c->opt.coarse_nodes : 0;
    if (want == c->coarse_built_for && (want == 0 || c->coarse_pat == c->vals.p)) return;
    if (c->side) HIPX(hipStreamSynchronize(c->side));  // no inversion in flight on the old buffers
    c->refresh_pending = false;
    c->coarse_built_for = 0;
    c->coarse_ready = false;
    if (want == 0) return;
    CoarseSpace C;
    std::string msg;
    int target = coarse_target(c->N, want);
    for (;;) {
        if (!build_coarse(c->N, c->h_rest.data(), c->has_mask ? c->h_mask.data() : nullptr, c->pat, target, C, &msg))
            throw HipError{"coarse space: " + msg};
        if (6 * C.nagg <= kMaxCoarse) break;
        target *= 2;  // keep the dense coarse inverse bounded
    }
    if (C.nagg == 0) {
        // every node fixed: nothing to correct, block-Jacobi alone
        c->nagg = c->nc = c->npad = 0;
        c->coarse_built_for = want;
        c->coarse_pat = c->vals.p;
        return;
    }
    c->nagg = C.nagg;
    c->nc = 6 * C.nagg;
    c->npairs = (int)C.pairI.size();
    c->ngchunks = (int)C.gptr.size() - 1;
    hipStream_t st = c->stream;
    auto up = [&](DevMem& m, const auto& v) { upload_vec(m, v, st); };
    up(c->cdn, C.dn);
    up(c->canodes, C.anodes);
    up(c->captr, C.aptr);
    up(c->crot, C.rot);
    up(c->cpairI, C.pairI);
    up(c->cpairJ, C.pairJ);
    up(c->cpchunk, C.pchunk);
    up(c->cgptr, C.gptr);
    up(c->cci, C.ci);
    up(c->ccj, C.cj);
    up(c->ccbase, C.cbase);
    const size_t nc2 = (size_t)std::max(c->nc, 1) * std::max(c->nc, 1);
    c->npad = (std::max(c->nc, 1) + kGjTile - 1) / kGjTile * kGjTile;
    const size_t np2 = (size_t)c->npad * c->npad;
    ensure(c->cE, sizeof(double) * nc2);
    ensure(c->cEinv, sizeof(double) * nc2);
    ensure(c->cEinv32, sizeof(float) * nc2);
    ensure(c->cX, sizeof(double) * np2);
    ensure(c->cY, sizeof(double) * np2);
    ensure(c->cpinv, sizeof(double) * 2 * kGjPivot * kGjPivot);
    ensure(c->crc, sizeof(double) * std::max(c->nc, 1));
    ensure(c->cgpart, sizeof(double) * 36 * std::max(c->ngchunks, 1));
    ensure(c->cpart, sizeof(double) * 2 * kMaxRedBlocks);
    ensure(c->cinfo, sizeof(int));
    if (!c->side) {
        HIPX(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
        HIPX(hipEventCreateWithFlags(&c->evG, hipEventDisableTiming));
        HIPX(hipEventCreateWithFlags(&c->evR, hipEventDisableTiming));
    }
    sync(c);
    c->coarse_built_for = want;
    c->coarse_pat = c->vals.p;
    c->coarse_tangents = 0;
    ++c->coarse_gen;
}

// the coarse inverse runs on the context's second stream (default priority):
// E padded, then npad / kGjPivot Gauss-Jordan steps; E^-1 in c->cres and the
// pivot check in cinfo when it ends (evR)
void coarse_factor_async(sors_ctx* c) {
    const int np = c->npad;
    const long long np2 = (long long)np * np;
    HIPX(hipEventRecord(c->evG, c->stream));  // E assembled
    HIPX(hipStreamWaitEvent(c->side, c->evG, 0));
    fill<int>(c->cinfo.as<int>(), 1, 0, c->side);
    k_coarse_pad<<<(unsigned)((np2 + kBlock - 1) / kBlock), kBlock, 0, c->side>>>(c->nc, np, c->cE.as<double>(),
                                                                                   c->cX.as<double>());
    double* in = c->cX.as<double>();
    double* out = c->cY.as<double>();
    double* pinv = c->cpinv.as<double>();  // two 32 x 32 inverses, by step parity
    const unsigned tiles = (unsigned)(np / kGjTile);
    k_gj_pivot<<<1, kGjThreads, 0, c->side>>>(np, in, pinv, c->cinfo.as<int>());
    for (int k0 = 0, s = 0; k0 < np; k0 += kGjPivot, ++s) {
        double* pc = pinv + (s & 1) * kGjPivot * kGjPivot;
        double* pn = pinv + ((s + 1) & 1) * kGjPivot * kGjPivot;
        k_spd_gj_step<<<tiles * tiles + 1, kGjThreads, 0, c->side>>>(np, k0, in, pc, pn, out, c->cinfo.as<int>());
        std::swap(in, out);
    }
    c->cres = in;
    HIPX(hipGetLastError());
    HIPX(hipEventRecord(c->evR, c->side));
    c->refresh_pending = true;
}

// E^-1 in a solve's precision: the FP32 copy for FP32 solves.  The additive
// two-level preconditioner D^-1 + Z E^-1 Z^T stays SPD with E^-1 rounded to
// FP32 (the rounding is far below the block-Jacobi term), and the apply
// still accumulates in FP64
template <typename HT>
const HT* coarse_inverse(const sors_ctx* c) {
    if constexpr (sizeof(HT) == 4)
        return c->cEinv32.as<float>();
    else
        return c->cEinv.as<double>();
}

// wait for the inversion in flight; on success the main stream installs
// E^-1 = (X + X^T)/2 and the coarse correction is on
void coarse_collect(sors_ctx* c) {
    if (!c->refresh_pending) return;
    c->refresh_pending = false;
    int h_info = 1;
    HIPX(hipMemcpyAsync(&h_info, c->cinfo.p, sizeof(int), hipMemcpyDeviceToHost, c->side));
    HIPX(hipStreamSynchronize(c->side));
    if (h_info != 0) return;  // E not positive definite: the previous inverse (if any) stays
    HIPX(hipStreamWaitEvent(c->stream, c->evR, 0));
    const size_t nc2 = (size_t)c->nc * c->nc;
    k_coarse_symmetrize<<<(unsigned)((nc2 + kBlock - 1) / kBlock), kBlock, 0, c->stream>>>(
        c->nc, c->npad, c->cres, c->cEinv.as<double>(), c->cEinv32.as<float>());
    HIPX(hipGetLastError());
    c->coarse_ready = true;
}

// E = Z^T A Z and its inverse (blocked Gauss-Jordan).  The first refresh,
// and every refresh when coarse_every <= 1, completes before the solve.
// Later refreshes (coarse_every > 1) invert on the second stream while the
// solve runs with the previous inverse; the next tangent installs the new
// one.
template <typename HT>
void coarse_setup(sors_ctx* c) {
    ensure_coarse(c);
    if (c->coarse_built_for == 0 || c->nc == 0) {
        c->coarse_ready = false;
        return;
    }
    coarse_collect(c);
    const long long k = c->coarse_tangents++;
    const bool lagged = c->opt.coarse_every > 1;
    if (c->coarse_ready && lagged && k % c->opt.coarse_every != 0) return;  // lagged operator
    const int nc = c->nc;
    const size_t nc2 = (size_t)nc * nc;
    fill<double>(c->cE.as<double>(), (long long)nc2, 0.0, c->stream);
    if (c->ngchunks > 0) {
        k_coarse_galerkin<HT><<<(c->ngchunks * 64 + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(
            c->ngchunks, c->cgptr.as<int>(), c->cci.as<int>(), c->ccj.as<int>(), c->ccbase.as<int>(),
            (const float4*)c->cdn.p, (const HT*)c->vals.p, c->cgpart.as<double>());
        k_coarse_gather<<<grid_for(36LL * c->npairs), kBlock, 0, c->stream>>>(
            c->npairs, nc, c->cpairI.as<int>(), c->cpairJ.as<int>(), c->cpchunk.as<int>(), c->crot.as<int>(),
            c->cgpart.as<double>(), c->cE.as<double>());
    }
    HIPX(hipGetLastError());
    coarse_factor_async(c);
    if (c->coarse_ready && lagged) return;  // the solve keeps the previous inverse
    c->coarse_ready = false;
    coarse_collect(c);  // the solve waits for this one
}

template <typename HT>
void build_preconditioner(sors_ctx* c) {
    if (c->block())
        k_block_jacobi_blk<HT><<<grid_for(c->N), kBlock, 0, c->stream>>>(
            c->N, c->sbase.as<int>(), c->selfslot.as<int>(), (const HT*)c->vals.p, (HT*)c->Minv.p);
    else
        k_block_jacobi<HT><<<grid_for(c->N), kBlock, 0, c->stream>>>(c->N, c->rowptr.as<int>(), c->cols.as<int>(),
                                                                      c->diag.as<int>(), (const HT*)c->vals.p,
                                                                      (HT*)c->Minv.p);
    HIPX(hipGetLastError());
    coarse_setup<HT>(c);
}

// launch geometry of the PCG kernels
#ifndef SORS_SPMV_BLOCKS
#define SORS_SPMV_BLOCKS 2048
#endif
#ifndef SORS_UPD_BLOCKS
#define SORS_UPD_BLOCKS 512
#endif
constexpr int kSpmvBlocks = SORS_SPMV_BLOCKS;   // multiple of 8 (XCD-contiguous row ranges)
#ifndef SORS_UPD_MIN_PER_THREAD
#define SORS_UPD_MIN_PER_THREAD 4
#endif
constexpr int kUpdMinPerThread = SORS_UPD_MIN_PER_THREAD;
constexpr int kUpdBlocks = SORS_UPD_BLOCKS;
static_assert(kSpmvBlocks <= kMaxRedBlocks && kUpdBlocks <= kMaxRedBlocks,
              "a PCG kernel's blocks each write a partial-sum slot (kMaxRedBlocks per reduction)");

struct PcgGrid {
    int gn;     // init (per node)
    int gu;     // DOF-parallel update
    int gs;     // SpMV
};

PcgGrid pcg_grid(sors_ctx* c) {
    PcgGrid g;
    g.gn = std::min(red_grid(c->N, kBlock), kUpdBlocks);
    // at least kUpdMinPerThread DOFs per thread: every workgroup also pays the
    // prologue that sums the previous kernels' partials
    g.gu = std::min(red_grid(c->n, kUpdBlock * kUpdMinPerThread), kUpdBlocks);
    long long work = c->block() ? (c->pat.nslices + kBlock / kSlice - 1) / (kBlock / kSlice)
                                : (c->n + kBlock / 8 - 1) / (kBlock / 8);
    g.gs = (int)std::min<long long>(std::max<long long>(work, 1), kSpmvBlocks);
    g.gs = (g.gs + kNumXcd - 1) / kNumXcd * kNumXcd;
    return g;
}

template <typename HT>
PcgView<HT> pcg_view(sors_ctx* c, double rtol) {
    PcgView<HT> v;
    v.N = c->N;
    v.n = c->n;
    v.nslices = c->pat.nslices;
    v.sbase = c->sbase.as<int>();
    v.bcol = c->bcol.as<int>();
    if (c->col16) v.bcol16 = c->bcol16.as<short>();
    v.vals = (const HT*)c->vals.p;
    v.Minv = (const HT*)c->Minv.p;
    v.x = (HT*)c->px.p;
    v.r = (HT*)c->pr.p;
    v.u = (HT*)c->pu.p;
    v.p = (HT*)c->pp.p;
    v.s = (HT*)c->ps.p;
    v.w = (HT*)c->pw.p;
    v.ctl = c->pcgctl.as<PcgCtl>();
    v.st = c->pcgst.as<PcgState>();
    v.rtol = rtol;
    return v;
}

// A_k (first: A_0) reading partU[par]; nbU = blocks of the producer of partU
// (with e0/e1: launched through hipExtLaunchKernelGGL, which records the two
// events around this dispatch itself)
template <typename HT>
void launch_spmv(sors_ctx* c, const PcgGrid& g, const PcgView<HT>& v, int par, int nbU, int first,
                 hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr, const double* inC = nullptr, int nbC = 0) {
    const PcgPart part = c->pcg_part();
    const PartIO io{part.a, kMaxRedBlocks, part.u[par], nbU, nullptr, 0, inC, nbC};
    PcgState* st = v.st + par;
    if (c->block() && e0)
        hipExtLaunchKernelGGL(k_pcg_spmv_blk<HT>, dim3(g.gs), dim3(kBlock), 0, c->stream, e0, e1, 0, v, st, io,
                              first);
    else if (c->block())
        k_pcg_spmv_blk<HT><<<g.gs, kBlock, 0, c->stream>>>(v, st, io, first);
    else if (e0)
        hipExtLaunchKernelGGL(k_pcg_spmv<HT, 8>, dim3(g.gs), dim3(kBlock), 0, c->stream, e0, e1, 0, c->n,
                              (const int*)c->rowptr.as<int>(), (const int*)c->cols.as<int>(), (const HT*)c->vals.p,
                              (const HT*)v.u, v.w, v.ctl, st, io, first, v.rtol);
    else
        k_pcg_spmv<HT, 8><<<g.gs, kBlock, 0, c->stream>>>(c->n, c->rowptr.as<int>(), c->cols.as<int>(),
                                                          (const HT*)c->vals.p, v.u, v.w, v.ctl, st, io, first,
                                                          v.rtol);
}

// U_k reading partU[par ^ 1] (produced with nbU blocks) and partA; writes
// partU[par] and st[par]
template <typename HT>
void launch_update(sors_ctx* c, const PcgGrid& g, const PcgView<HT>& v, int par, int nbU, int first,
                   const double* inC = nullptr, int nbC = 0, const PcgProj<HT>* proj = nullptr) {
    const PcgPart part = c->pcg_part();
    const PartIO io{part.u[par], kMaxRedBlocks, part.u[par ^ 1], nbU, part.a, g.gs, inC, nbC};
    if (proj)
        k_pcg_update_proj<HT><<<g.gu, kUpdBlock, 0, c->stream>>>(v, v.st + (par ^ 1), v.st + par, io, first, *proj);
    else
        k_pcg_update<HT><<<g.gu, kUpdBlock, 0, c->stream>>>(v, v.st + (par ^ 1), v.st + par, io, first);
}

// u += Z E^-1 Z^T r after an update (or the init); gamma's coarse partials into slot par
template <typename HT>
void launch_coarse(sors_ctx* c, PcgCtl* ctl, int par) {
    const CoarseArgs ca = coarse_args(c);
    k_coarse_restrict<HT><<<c->nagg, kRestrictThreads, 0, c->stream>>>(ca, ctl, (const HT*)c->pr.p);
    k_coarse_apply<HT><<<c->nagg, kApplyThreads, 0, c->stream>>>(ca, ctl, coarse_inverse<HT>(c), (HT*)c->pu.p,
                                                                 c->cpart.as<double>() + par * kMaxRedBlocks);
}

// PCG on the assembled matrix: solves A x = sign*b; x in c->px (HT).
// Sequence: init, A_0, U_1, A_1, then chunks of an even number of
// iterations (U_k, A_k), replayed from a captured hipGraph when use_graph.
// The host reads the outcome record between chunks.  With profile_kernels
// the first chunk of every solve is launched directly with a HIP event pair
// around each SpMV.  With proj the residual is projected on the constrained
// nodes (minimize_SQP; launched directly, without the coarse correction).
template <typename HT>
long long pcg(sors_ctx* c, const double* b, double sign, double rtol, int max_iter, bool* breakdown,
              const PcgProj<HT>* proj = nullptr) {
    const PcgView<HT> v = pcg_view<HT>(c, rtol);
    PcgCtl* ctl = v.ctl;
    const PcgGrid g = pcg_grid(c);
    const int chunk = std::max(2, (c->opt.pcg_check_every + 1) / 2 * 2);
    const int gi = std::is_same<HT, float>::value ? 1 : 0;
    hipEvent_t* ev = c->prof ? c->events(2 * (chunk + 1)) : nullptr;
    auto account = [&](int count) {
        for (int j = 0; j < count; ++j) {
            float ms = 0;
            HIPX(hipEventElapsedTime(&ms, ev[2 * j], ev[2 * j + 1]));
            c->prof->spmv_ms += ms;
            c->prof->spmv_launches += 1;
        }
    };
    // two-level preconditioner: a coarse correction after every update
    const bool cs = c->coarse_ready && !proj;
    const int nbC = cs ? c->nagg : 0;
    auto inC = [&](int par) -> const double* { return cs ? c->cpart.as<double>() + par * kMaxRedBlocks : nullptr; };
    // iteration k >= 2 reads from st[(k-1)%2] and writes st[k%2]
    auto iteration = [&](int k, int j, bool timed) {
        const int par = k & 1;
        launch_update<HT>(c, g, v, par, g.gu, 0, inC(par ^ 1), nbC, proj);
        if (cs) launch_coarse<HT>(c, ctl, par);
        if (timed)
            launch_spmv<HT>(c, g, v, par, g.gu, 0, ev[2 * j], ev[2 * j + 1], inC(par), nbC);
        else
            launch_spmv<HT>(c, g, v, par, g.gu, 0, nullptr, nullptr, inC(par), nbC);
    };
    PcgCtl* hctl = (PcgCtl*)c->h_pin;
    fill<int>((int*)ctl, sizeof(PcgCtl) / sizeof(int), 0, c->stream);
    {
        const PcgPart part = c->pcg_part();
        const PartIO io{part.u[0], kMaxRedBlocks, nullptr, 0, nullptr, 0};
        if (proj)
            k_pcg_init_proj<HT><<<g.gn, kBlock, 0, c->stream>>>(v, b, sign, io, *proj);
        else
            k_pcg_init<HT><<<g.gn, kBlock, 0, c->stream>>>(v, b, sign, io);
    }
    if (cs) launch_coarse<HT>(c, ctl, 0);
    launch_spmv<HT>(c, g, v, 0, g.gn, 1, nullptr, nullptr, inC(0), nbC);   // A_0
    launch_update<HT>(c, g, v, 1, g.gn, 1, inC(0), nbC, proj);              // U_1
    if (cs) launch_coarse<HT>(c, ctl, 1);
    launch_spmv<HT>(c, g, v, 1, g.gu, 0, nullptr, nullptr, inC(1), nbC);   // A_1
    HIPX(hipGetLastError());
    // chunk sizes (even: the parity pattern repeats); after the first chunk
    // the next one is the smallest size that should finish the solve at the
    // convergence rate seen so far, so a converged solve leaves few launches
    // that return at once
    const int sizes[3] = {std::max(2, chunk / 8 * 2), std::max(2, chunk / 4 * 2), chunk};
    PcgState* hst = (PcgState*)(c->h_pin + 8);  // after the PcgCtl record
    auto next_size = [&]() {
        const PcgState& st = hst[0].iter >= hst[1].iter ? hst[0] : hst[1];
        if (!(st.iter > 1 && st.rr0 > 0.0 && st.rr > st.tol2 && st.rr < st.rr0 && st.tol2 > 0.0)) return chunk;
        // st.rr belongs to iteration st.iter - 1; two iterations of margin
        const double rate = std::log(st.rr / st.rr0) / (st.iter - 1);
        const double left = std::log(st.tol2 / st.rr) / rate - 1.0 + 2.0;
        for (int sz : sizes)
            if (sz >= left) return sz;
        return chunk;
    };
    int launched = 1, next = chunk;
    bool first_chunk = true, done = false;
    while (launched < max_iter) {
        const int k = std::min(next, max_iter - launched);
        const bool timed = ev && first_chunk;
        int si = -1;
        for (int q = 0; q < 3; ++q)
            if (k == sizes[q]) si = q;
        if (!timed && c->opt.use_graph && si >= 0 && !proj) {
            hipGraphExec_t& ge = c->gexec[gi][si];
            if (!ge || c->gkey[gi][si] != c->vals.p || c->gchunk[gi][si] != k ||
                c->gcoarse[gi][si] != (cs ? c->coarse_gen : 0)) {
                if (ge) HIPX(hipGraphExecDestroy(ge));
                ge = nullptr;
                hipGraph_t graph = nullptr;
                HIPX(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                for (int j = 0; j < k; ++j) iteration(launched + 1 + j, j, false);  // parity pattern repeats
                HIPX(hipStreamEndCapture(c->stream, &graph));
                HIPX(hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0));
                HIPX(hipGraphDestroy(graph));
                c->gkey[gi][si] = c->vals.p;
                c->gchunk[gi][si] = k;
                c->gcoarse[gi][si] = cs ? c->coarse_gen : 0;
            }
            HIPX(hipGraphLaunch(ge, c->stream));
        } else {
            for (int j = 0; j < k; ++j) iteration(launched + 1 + j, j, timed);
        }
        HIPX(hipGetLastError());
        launched += k;
        HIPX(hipMemcpyAsync(hctl, ctl, sizeof(PcgCtl), hipMemcpyDeviceToHost, c->stream));
        HIPX(hipMemcpyAsync(hst, v.st, 2 * sizeof(PcgState), hipMemcpyDeviceToHost, c->stream));
        sync(c);
        if (timed) {
            // SpMV launches after the one that ended the solve return at once: not counted
            const int worked =
                hctl->done ? std::max(0, hctl->iter - (launched - k) - (hctl->breakdown ? 0 : 1)) : k;
            account(std::min(worked, k));
        }
        first_chunk = false;
        if (hctl->done) {
            done = true;
            break;
        }
        next = next_size();
    }
    if (!done) {
        HIPX(hipMemcpyAsync(hctl, ctl, sizeof(PcgCtl), hipMemcpyDeviceToHost, c->stream));
        sync(c);
    }
    if (breakdown) *breakdown = hctl->breakdown != 0;
    return hctl->done ? hctl->iter : launched;
}

double read_scalar(sors_ctx* c, int slot) {
    HIPX(hipMemcpyAsync(c->h_pin, c->S() + slot, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    sync(c);
    return c->h_pin[0];
}

// ||g||^2 (scal[S_GG]) and the inverted-element flag of an assembly, read
// back with one wait
double read_gg_flag(sors_ctx* c, int* flag) {
    HIPX(hipMemcpyAsync(c->h_pin, c->S() + S_GG, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPX(hipMemcpyAsync(c->h_pin + 1, c->F(), sizeof(int), hipMemcpyDeviceToHost, c->stream));
    sync(c);
    *flag = *(const int*)(c->h_pin + 1);
    return c->h_pin[0];
}

void read_scalars(sors_ctx* c, double* out, int first, int count) {
    HIPX(hipMemcpyAsync(c->h_pin, c->S() + first, sizeof(double) * count, hipMemcpyDeviceToHost,
                        c->stream));
    sync(c);
    for (int i = 0; i < count; ++i) out[i] = c->h_pin[i];
}

// Where the energy kernels put their work: the stream, the reduction
// partials, the three result slots (element, node, contact) and the
// inverted-element flag.  The context's own buffers by default; the
// allocation-free entry point passes the device's static scratch.
struct EnergyTarget {
    hipStream_t st;
    double* partials;
    double* out3;
    int* flag;
};

EnergyTarget ctx_energy_target(sors_ctx* c) {
    return EnergyTarget{c->stream, c->partials.as<double>(), c->S() + S_EELEM, c->F()};
}

// energy kernels into out3[0] (elements), out3[1] (node terms) and out3[2]
// (contact): with dx, the line-search difference E(q + alpha dx) - E(q);
// without, E(q) (FP64, device)
void enqueue_energy(sors_ctx* c, const double* q, const double* dx, double alpha, double dt, const EnergyTarget& t) {
    const int diff = dx != nullptr;
    const Reduce red{t.partials};
    ElemArgs a = elem_args(c, q);
    a.dq = dx;
    a.alpha = alpha;
    a.flag = t.flag;
    const int ge = red_grid(c->E, kBlock);
    if (c->ed == 4)
        k_elem_energy<4><<<ge, kBlock, 0, t.st>>>(a, diff, t.out3, red);
    else
        k_elem_energy<8><<<ge, kBlock, 0, t.st>>>(a, diff, t.out3, red);
    k_finalize<1><<<1, kBlock, 0, t.st>>>(t.partials, ge, t.out3);
    NodeArgs na = node_args(c, q, dt);
    if (c->has_pressure) {
        pressure_force(c, q, nullptr, 0.0, c->fp0.as<double>(), t.st);
        if (diff) {
            pressure_force(c, q, dx, alpha, c->fp1.as<double>(), t.st);
            na.fpres1 = c->fp1.as<double>();
        }
    }
    const int gnn = red_grid(c->n, kBlock);
    k_node_energy<<<gnn, kBlock, 0, t.st>>>(na, dx, alpha, diff, t.out3 + 1, red);
    k_finalize<1><<<1, kBlock, 0, t.st>>>(t.partials, gnn, t.out3 + 1);
    if (penalty_on(c)) {
        const int gcn = red_grid(c->cargs.nsurf, kBlock);
        k_contact_energy<<<gcn, kBlock, 0, t.st>>>(contact_args(c), q, dx, alpha, diff, c->qprev.as<double>(), dt,
                                                   red);
        k_finalize<1><<<1, kBlock, 0, t.st>>>(t.partials, gcn, t.out3 + 2);
    }
    HIPX(hipGetLastError());
}

void enqueue_energy(sors_ctx* c, const double* q, const double* dx, double alpha, double dt) {
    enqueue_energy(c, q, dx, alpha, dt, ctx_energy_target(c));
}

// the energy scalars summed on the host: element + node (+ contact)
double energy_sum(const sors_ctx* c, const double* s3) {
    return s3[0] + s3[1] + (penalty_on(c) ? s3[2] : 0.0);
}

static_assert(S_ENODE == S_EELEM + 1 && S_AUX == S_EELEM + 2, "energy scalars are read as one triple");

double energy_delta(sors_ctx* c, const double* q, const double* dx, double alpha, double dt) {
    enqueue_energy(c, q, dx, alpha, dt);
    double v[3];
    read_scalars(c, v, S_EELEM, 3);
    return energy_sum(c, v);
}

double energy_abs(sors_ctx* c, const double* q, double dt) {
    enqueue_energy(c, q, nullptr, 0.0, dt);
    double v[3];
    read_scalars(c, v, S_EELEM, 3);
    return energy_sum(c, v);
}

// step dots into S[S_DD], S[S_GD] and the step dx = x (FP64)
template <typename HT>
void enqueue_step_dots(sors_ctx* c, const double* g, double* dx) {
    k_to_f64<HT><<<red_grid(c->n, kBlock), kBlock, 0, c->stream>>>(c->n, (const HT*)c->px.p, dx);
    const int gd = red_grid(c->n, kBlock);
    k_step_dots<HT><<<gd, kBlock, 0, c->stream>>>(c->n, (const HT*)c->px.p, g, c->S() + S_DD, c->S() + S_GD,
                                                   c->red());
    static_assert(S_GD == S_DD + 1, "step dots are finalized as one pair");
    k_finalize<2><<<1, kBlock, 0, c->stream>>>(c->partials.as<double>(), gd, c->S() + S_DD);
    HIPX(hipGetLastError());
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Solver::minimize_frozenhessian (solver.cpp:268-294): the tangent at the
// initial guess is assembled, and its preconditioner built, once; every
// iteration assembles the gradient alone and solves with that tangent; no
// line search.  max_newton iterations without ||dx|| < dx_tol end the step
// without an error, as the reference returns its iterate.
template <typename HT>
int minimize_frozen(sors_ctx* c, double dt, sors_stats* stt) {
    const sors_options& o = c->opt;
    double* q = c->q.as<double>();
    double* g = c->g.as<double>();
    double* dx = c->dxd.as<double>();
    for (int it = 0; it < o.max_newton; ++it) {
        auto t0 = std::chrono::steady_clock::now();
        clear_flag(c);
        assemble<HT>(c, q, dt, it == 0, o.project_spd);
        int flag = 0;
        const double gg = read_gg_flag(c, &flag);
        if (flag & 1) return fail(c, SORS_ERR_INVERTED, "Inverted Elements");
        stt->gnorm = std::sqrt(gg);
        if (!(gg == gg)) return fail(c, SORS_ERR_NAN, "Frozen Newton's method diverged (NaN residual)");
        if (o.g_tol > 0 && stt->gnorm < o.g_tol) {
            stt->ms_assemble += ms_since(t0);
            break;
        }
        if (it == 0) build_preconditioner<HT>(c);
        stt->ms_assemble += ms_since(t0);
        auto t1 = std::chrono::steady_clock::now();
        bool brk = false;
        stt->pcg_iters += pcg<HT>(c, g, -1.0, o.pcg_rtol, o.pcg_max_iter, &brk);
        enqueue_step_dots<HT>(c, g, dx);
        double dots[2];
        read_scalars(c, dots, S_DD, 2);
        stt->ms_solve += ms_since(t1);
        stt->newton_iters++;
        const double dxn = std::sqrt(dots[0]);
        if (!(dxn == dxn)) return fail(c, SORS_ERR_NAN, "Frozen Newton's method diverged");
        if (brk && dxn == 0.0) return fail(c, SORS_ERR_NAN, "PCG breakdown (indefinite tangent)");
        k_axpy<<<red_grid(c->n, kBlock), kBlock, 0, c->stream>>>(c->n, 1.0, dx, q);
        HIPX(hipGetLastError());
        stt->dxnorm = dxn;
        if (dxn < o.dx_tol) break;
    }
    return SORS_OK;
}

// relative PCG tolerance of Newton iteration `it` before the g_tol
// loosening: pcg_rtol_first (when set above pcg_rtol) for the first solve of
// a minimisation, whose reduction the nonlinearity of the first Newton step
// caps anyway, pcg_rtol for the others
// a solve aimed at the residual criterion brings its linear residual under
// kAim * g_tol, leaving the rest of g_tol to the Newton step's nonlinear
// remainder (0.7 and 0.8 were measured too: a few percent either way across
// the beam, hex and contact rollouts, and more failed residual checks)
constexpr double kAim = 0.5;

double forcing_base(const sors_options& o, int it) {
    return it == 0 && o.pcg_rtol_first > o.pcg_rtol ? o.pcg_rtol_first : o.pcg_rtol;
}

// a residual of norm g is within reach of g_tol when the solve that follows
// would be aimed at it (its forcing term would come from g_tol)
bool in_reach(const sors_options& o, int it, double g) {
    return o.g_tol > 0 && g > 0 && kAim * o.g_tol / g > forcing_base(o, it);
}

// options.reuse_tangent: Newton iteration `it` evaluates its gradient alone
// first and, should the residual be within reach of g_tol, solves with the
// tangent (and preconditioner) of the last solve once more, when that last
// solve's forcing term eta applied to its residual g predicts a residual
// within reach and the last solve did not reuse its tangent already
bool may_reuse_tangent(const sors_options& o, int it, bool reused, double eta, double g) {
    return o.reuse_tangent && it >= 1 && !reused && in_reach(o, it, eta * g);
}

// minimize_newton / minimize_newton_linesearch on the device state c->q
// (minimize_frozenhessian above)
template <typename HT>
int minimize(sors_ctx* c, double dt, sors_stats* stt) {
    if (c->method == SORS_MINIMIZE_FROZENHESSIAN) return minimize_frozen<HT>(c, dt, stt);
    const sors_options& o = c->opt;
    double* q = c->q.as<double>();
    double* g = c->g.as<double>();
    double* dx = c->dxd.as<double>();
    bool converged = false;
    // the previous solve was aimed at the residual criterion (its forcing
    // term came from g_tol): the next residual most likely passes, so its
    // gradient is assembled alone first and the tangent only if it does not
    bool aimed = false, reused = false;
    double last_eta = 0.0;
    for (int it = 0; it < o.max_newton; ++it) {
        auto t0 = std::chrono::steady_clock::now();
        bool assembled = false, reuse = false;
        const bool may_reuse = may_reuse_tangent(o, it, reused, last_eta, stt->gnorm);
        for (int pass = aimed || may_reuse ? 0 : 1; pass < 2 && !assembled; ++pass) {
            clear_flag(c);
            assemble<HT>(c, q, dt, pass == 1, o.project_spd);
            int flag = 0;
            const double gg = read_gg_flag(c, &flag);
            if (flag & 1) return fail(c, SORS_ERR_INVERTED, "Inverted Elements");
            stt->gnorm = std::sqrt(gg);
            if (!(gg == gg)) return fail(c, SORS_ERR_NAN, "Newton's method diverged (NaN residual)");
            if (o.g_tol > 0 && stt->gnorm < o.g_tol) {
                converged = true;
                break;
            }
            if (pass == 0 && may_reuse && in_reach(o, it, stt->gnorm)) {
                reuse = true;  // the matrix and preconditioner of the last solve
                break;
            }
            assembled = pass == 1;
        }
        if (converged) {
            stt->ms_assemble += ms_since(t0);
            break;
        }
        if (!reuse) build_preconditioner<HT>(c);  // only for a tangent that is solved with
        reused = reuse;
        stt->ms_assemble += ms_since(t0);
        auto t1 = std::chrono::steady_clock::now();
        bool brk = false;
        // forcing term: pcg_rtol (pcg_rtol_first for the minimisation's first
        // solve), loosened when the residual criterion is in reach (the linear
        // residual of the step only has to drop below g_tol: eta ||g|| <
        // g_tol / 2), never above 0.1
        const double base = forcing_base(o, it);
        double eta = base;
        if (o.g_tol > 0 && stt->gnorm > 0) eta = std::min(0.1, std::max(eta, kAim * o.g_tol / stt->gnorm));
        aimed = o.g_tol > 0 && eta > base;
        last_eta = eta;
        stt->pcg_iters += pcg<HT>(c, g, -1.0, eta, o.pcg_max_iter, &brk);
        enqueue_step_dots<HT>(c, g, dx);
        // Newton + Armijo with a residual criterion: the line search's first
        // trial (alpha = 1) is evaluated with the step dots and read back with
        // them (one wait fewer per iteration)
        static_assert(S_GD == S_DD + 1 && S_EELEM == S_DD + 2, "dots and energy scalars are read as one range");
        const bool first_trial = c->method == SORS_MINIMIZE_NEWTON_LINESEARCH && o.g_tol > 0;
        if (first_trial) enqueue_energy(c, q, dx, 1.0, dt);
        double dots[5];
        read_scalars(c, dots, S_DD, first_trial ? 5 : 2);
        stt->ms_solve += ms_since(t1);
        stt->newton_iters++;
        const double dxn = std::sqrt(dots[0]), gdx = dots[1];
        if (!(dxn == dxn)) return fail(c, SORS_ERR_NAN, "Newton's method diverged");
        if (brk && dxn == 0.0) return fail(c, SORS_ERR_NAN, "PCG breakdown (indefinite tangent)");
        if (c->method != SORS_MINIMIZE_NEWTON_LINESEARCH || dxn < o.dx_tol) {
            k_axpy<<<red_grid(c->n, kBlock), kBlock, 0, c->stream>>>(c->n, 1.0, dx, q);
            stt->dxnorm = dxn;
            if (dxn < o.dx_tol) {
                converged = true;
                break;
            }
            continue;
        }
        // Armijo backtracking (solver.cpp:225-240), energy decrease in FP64
        auto t2 = std::chrono::steady_clock::now();
        double alpha = 1.0;
        bool accepted = false;
        for (int bt = 0; bt <= o.max_backtrack; ++bt) {
            const double dE = bt == 0 && first_trial ? energy_sum(c, dots + 2) : energy_delta(c, q, dx, alpha, dt);
            if (dE <= 0.1 * alpha * gdx) {
                accepted = true;
                break;
            }
            alpha *= 0.5;
            stt->backtracks++;
        }
        stt->ms_linesearch += ms_since(t2);
        if (!accepted) {
            if (alpha * 2 * dxn < o.dx_tol)
                alpha *= 2;
            else
                return fail(c, SORS_ERR_LINESEARCH, "line search failed to decrease the energy");
        }
        k_axpy<<<red_grid(c->n, kBlock), kBlock, 0, c->stream>>>(c->n, alpha, dx, q);
        HIPX(hipGetLastError());
        stt->dxnorm = alpha * dxn;
        if (stt->dxnorm < o.dx_tol) {
            converged = true;
            break;
        }
    }
    if (!converged) return fail(c, SORS_ERR_NOCONV, "Newton's method did not converge");
    return SORS_OK;
}

// ---- minimize_SQP -------------------------------------------------------
constexpr int kSqpMaxWorkingSet = 40;  // working-set iterations per QP

// the K table (nodes with Dirichlet DOFs or on the contact candidate list)
// and the SQP buffers; built once per contact / mask set-up
void ensure_sqp(sors_ctx* c) {
    if (c->sqp_built) return;
    const bool con = c->has_contact && c->cargs.nsurf > 0;
    std::vector<int> eqbits((size_t)c->N, 0);
    std::vector<char> isK((size_t)c->N, 0);
    if (c->has_mask)
        for (int i = 0; i < c->n; ++i)
            if (c->h_mask[(size_t)i]) {
                eqbits[(size_t)i / 3] |= 1 << (i % 3);
                isK[(size_t)i / 3] = 1;
            }
    if (con)
        for (int v : c->h_surf) isK[(size_t)v] = 1;
    std::vector<int> knode, keq, kmap((size_t)c->N, -1), surf2k;
    for (int v = 0; v < c->N; ++v)
        if (isK[(size_t)v]) {
            kmap[(size_t)v] = (int)knode.size();
            knode.push_back(v);
            keq.push_back(eqbits[(size_t)v]);
        }
    if (con)
        for (int v : c->h_surf) surf2k.push_back(kmap[(size_t)v]);
    c->nK = (int)knode.size();
    hipStream_t st = c->stream;
    upload_vec(c->sknode, knode, st);
    upload_vec(c->skeq, keq, st);
    upload_vec(c->skmap, kmap, st);
    upload_vec(c->ssurf2k, surf2k, st);
    const int nk1 = std::max(c->nK, 1), ns1 = std::max(con ? c->cargs.nsurf : 0, 1);
    ensure(c->skpid, sizeof(int) * nk1);
    ensure(c->skdid, sizeof(int) * nk1);
    ensure(c->skw, sizeof(int) * nk1);
    ensure(c->spmask, sizeof(int) * ns1);
    ensure(c->sdmask, sizeof(int) * ns1);
    ensure(c->sfrozen, sizeof(int));
    ensure(c->schg, sizeof(int));
    ensure(c->sP, sizeof(double) * 9 * nk1);
    ensure(c->sMsave, sizeof(double) * 9 * nk1);
    ensure(c->sdxfull, sizeof(double) * c->n);
    ensure(c->srhs, sizeof(double) * c->n);
    fill<int>(c->skpid.as<int>(), nk1, -1, st);
    fill<int>(c->skdid.as<int>(), nk1, -1, st);
    fill<int>(c->skw.as<int>(), nk1, 0, st);
    fill<int>(c->spmask.as<int>(), ns1, -1, st);
    fill<int>(c->sdmask.as<int>(), ns1, -1, st);
    fill<int>(c->sfrozen.as<int>(), 1, 0, st);
    // rows outside the working set count as violated beyond 1e-9 of the
    // rest bounding-box diagonal
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int v = 0; v < c->N; ++v)
        for (int k = 0; k < 3; ++k) {
            const double x = c->h_rest[3 * (size_t)v + k];
            if (v == 0 || x < lo[k]) lo[k] = x;
            if (v == 0 || x > hi[k]) hi[k] = x;
        }
    const double diag = std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                                  (hi[2] - lo[2]) * (hi[2] - lo[2]));
    c->sqp_vtol = 1e-9 * (diag > 0.0 ? diag : 1.0);
    sync(c);
    c->sqp_built = true;
}

SqpDev sqp_dev(sors_ctx* c) {
    SqpDev s{};
    s.nK = c->nK;
    s.nsurf = (c->has_contact && c->cargs.nsurf > 0) ? c->cargs.nsurf : 0;
    s.knode = c->sknode.as<int>();
    s.keq = c->skeq.as<int>();
    s.surf2k = c->ssurf2k.as<int>();
    s.kpid = c->skpid.as<int>();
    s.kdid = c->skdid.as<int>();
    s.kw = c->skw.as<int>();
    s.pmask = c->spmask.as<int>();
    s.dmask = c->sdmask.as<int>();
    s.frozen = c->sfrozen.as<int>();
    s.changes = c->schg.as<int>();
    s.dxfull = c->sdxfull.as<double>();
    s.vtol = c->sqp_vtol;
    return s;
}

// Solver::minimize_SQP (solver.cpp:109-197) on the device state c->q: the
// active sets from the initial guess (update_active: once per step, as the
// reference; an FP64 continuation keeps them), then up to max_newton SQP
// iterations -- gradient and tangent with the SQP semantics (no masking, no
// Dirichlet override, no penalty), the QP by working-set iterations of
// projected PCG solves, x += dx -- until ||dx|| < dx_tol.  No line search, as
// the reference.
template <typename HT>
int minimize_sqp(sors_ctx* c, double dt, sors_stats* stt, bool update_active) {
    if (!c->block()) return fail(c, SORS_ERR_UNSUPPORTED, "minimize_SQP needs the node-block layout");
    ensure_sqp(c);
    const sors_options& o = c->opt;
    hipStream_t st = c->stream;
    double* q = c->q.as<double>();
    double* g = c->g.as<double>();
    double* dx = c->dxd.as<double>();
    const SqpDev s = sqp_dev(c);
    const ContactArgs ca = contact_args(c);
    if (update_active && s.nsurf > 0) k_sqp_active<<<1, kSqpActiveThreads, 0, st>>>(s, ca, q);
    const PcgProj<HT> pj{c->skmap.as<int>(), (const HT*)c->sP.p};
    const int gK = grid_for(std::max(c->nK, 1));
    const double* qprev = c->qprev.as<double>();
    const double* bcval = c->bcval.as<double>();
    bool converged = false;
    for (int it = 0; it < o.max_newton; ++it) {
        auto t0 = std::chrono::steady_clock::now();
        clear_flag(c);
        assemble<HT>(c, q, dt, true, o.project_spd);
        int flag = 0;
        const double gg = read_gg_flag(c, &flag);
        if (flag & 1) return fail(c, SORS_ERR_INVERTED, "Inverted Elements");
        stt->gnorm = std::sqrt(gg);
        if (!(gg == gg)) return fail(c, SORS_ERR_NAN, "SQP method diverged (NaN residual)");
        build_preconditioner<HT>(c);
        if (c->nK) k_sqp_save_minv<HT><<<gK, kBlock, 0, st>>>(s, (const HT*)c->Minv.p, (HT*)c->sMsave.p);
        fill<double>(c->sdxfull.as<double>(), c->n, 0.0, st);
        stt->ms_assemble += ms_since(t0);
        auto t1 = std::chrono::steady_clock::now();
        bool settled = false;
        for (int pass = 0; pass < kSqpMaxWorkingSet; ++pass) {
            if (c->nK)
                k_sqp_setup<HT><<<gK, kBlock, 0, st>>>(s, ca, q, qprev, bcval, dt, (const HT*)c->sMsave.p,
                                                        (HT*)c->sP.p, (HT*)c->Minv.p);
            k_sqp_rhs<HT><<<grid_for(c->N), kBlock, 0, st>>>(c->N, c->sbase.as<int>(), c->bcol.as<int>(),
                                                              (const HT*)c->vals.p, c->sdxfull.as<double>(), g,
                                                              c->srhs.as<double>());
            HIPX(hipGetLastError());
            bool brk = false;
            stt->pcg_iters += pcg<HT>(c, c->srhs.as<double>(), 1.0, o.pcg_rtol, o.pcg_max_iter, &brk, &pj);
            stt->qp_iters++;
            if (!c->nK) {
                settled = true;
                break;
            }
            fill<int>(c->schg.as<int>(), 1, 0, st);
            k_sqp_multipliers<HT><<<gK, kBlock, 0, st>>>(s, ca, c->sbase.as<int>(), c->bcol.as<int>(),
                                                          (const HT*)c->vals.p, (const HT*)c->px.p, g, q, qprev,
                                                          bcval, dt);
            HIPX(hipGetLastError());
            int* hchg = (int*)c->h_pin;
            HIPX(hipMemcpyAsync(hchg, c->schg.p, sizeof(int), hipMemcpyDeviceToHost, st));
            sync(c);
            if (*hchg == 0) {
                settled = true;
                break;
            }
        }
        if (!settled) return fail(c, SORS_ERR_NOCONV, "SQP working set did not settle");
        k_sqp_step<HT><<<red_grid(c->n, kBlock), kBlock, 0, st>>>(c->n, c->sdxfull.as<double>(),
                                                                   (const HT*)c->px.p, dx);
        const int gd = red_grid(c->n, kBlock);
        k_sumsq<<<gd, kBlock, 0, st>>>(c->n, dx, c->red());
        k_finalize<1><<<1, kBlock, 0, st>>>(c->partials.as<double>(), gd, c->S() + S_DD);
        HIPX(hipGetLastError());
        const double dxn = std::sqrt(read_scalar(c, S_DD));
        stt->ms_solve += ms_since(t1);
        stt->newton_iters++;
        if (!(dxn == dxn)) return fail(c, SORS_ERR_NAN, "SQP method diverged");
        k_axpy<<<red_grid(c->n, kBlock), kBlock, 0, st>>>(c->n, 1.0, dx, q);
        HIPX(hipGetLastError());
        stt->dxnorm = dxn;
        if (dxn < o.dx_tol) {
            converged = true;
            break;
        }
    }
    if (!converged) return fail(c, SORS_ERR_NOCONV, "SQP method did not converge");
    return SORS_OK;
}

// Newton's starting point for a step of size h (options.predict: 0 keep q,
// 1 constant velocity, 2 constant acceleration once a previous step exists)
void launch_predict(sors_ctx* c, double h) {
    if (!c->opt.predict) return;
    const bool acc = c->opt.predict >= 2 && c->vold_valid && c->vold_dt > 0.0;
    k_predict<<<red_grid(c->n, kBlock), kBlock, 0, c->stream>>>(
        c->n, h, c->qprev.as<double>(), c->vprev.as<double>(), acc ? c->vold.as<double>() : nullptr,
        acc ? h / c->vold_dt : 0.0, c->has_mask ? c->mask.as<int>() : nullptr, c->bcval.as<double>(),
        c->q.as<double>());
}

// TimeIntegrator::update_state after a converged step of size h
void launch_update_state(sors_ctx* c, double h) {
    double* vold = nullptr;
    if (c->opt.predict >= 2) {
        ensure(c->vold, sizeof(double) * c->n);
        vold = c->vold.as<double>();
    }
    k_update_state<<<red_grid(c->n, kBlock), kBlock, 0, c->stream>>>(
        c->n, c->scheme, h, c->q.as<double>(), c->qprev.as<double>(), c->vprev.as<double>(), c->gprev.as<double>(),
        c->gtmp.as<double>(), vold);
    c->vold_valid = vold != nullptr;
    c->vold_dt = h;
}

// One minimisation of a step (the Newton family or SQP) with the mixed-
// precision continuations: a mixed solve that runs out of iterations goes on
// from its iterate with FP64 tangents and solves; one that inverts an
// element restarts from its initial guess in FP64 (an FP32 tangent or solve
// can overshoot where the FP64 ones do not).  SQP keeps the step's active sets.
int minimize_step(sors_ctx* c, double h, sors_stats* st) {
    const bool mixed = c->opt.precision == SORS_MIXED;
    if (mixed) {
        ensure(c->qtmp, sizeof(double) * c->n);
        HIPX(hipMemcpyAsync(c->qtmp.p, c->q.p, sizeof(double) * c->n, hipMemcpyDeviceToDevice, c->stream));
    }
    int rc;
    if (c->sqp())
        rc = mixed ? minimize_sqp<float>(c, h, st, true) : minimize_sqp<double>(c, h, st, true);
    else
        rc = mixed ? minimize<float>(c, h, st) : minimize<double>(c, h, st);
    if (mixed && (rc == SORS_ERR_NOCONV || rc == SORS_ERR_INVERTED)) {
        if (rc == SORS_ERR_INVERTED)
            HIPX(hipMemcpyAsync(c->q.p, c->qtmp.p, sizeof(double) * c->n, hipMemcpyDeviceToDevice, c->stream));
        st->fp64_fallbacks++;
        rc = c->sqp() ? minimize_sqp<double>(c, h, st, false) : minimize<double>(c, h, st);
    }
    return rc;
}

int do_steps(sors_ctx* c, double dt, int substeps, sors_stats* st) {
    if (!(dt > 0.0)) return fail(c, SORS_ERR_ARG, "dt must be > 0 (resolve settings.dt before the call)");
    if (substeps < 1) substeps = 1;
    sors_stats local{};
    if (!st) st = &local;
    *st = sors_stats{};
    c->prof = c->opt.profile_kernels ? st : nullptr;
    struct ProfReset {
        sors_ctx* c;
        ~ProfReset() { c->prof = nullptr; }
    } prof_reset{c};
    auto t0 = std::chrono::steady_clock::now();
    const double h = dt / substeps;
    ensure_pattern(c, h);
    for (int j = 0; j < substeps; ++j) {
        // minimize_SQP starts from the caller's guess (solver.cpp:30-31)
        if (!c->sqp()) {
            launch_predict(c, h);
            contact_lag(c);
        }
        const int rc = minimize_step(c, h, st);
        if (rc != SORS_OK) return rc;
        launch_update_state(c, h);
        HIPX(hipGetLastError());
        std::memcpy(c->cargs.dc_last, c->cargs.dc, sizeof(c->cargs.dc));
    }
    sync(c);
    st->ms_total = ms_since(t0);
    return SORS_OK;
}

// ===========================================================================
// Serving loop (no counterpart in the reference): continuous batching of
// independent rollouts.  Requests wait in arrival order; the head is
// admitted when a slot and HBM budget are free and handed to a prep thread
// (context, pattern, upload); resident rollouts advance in ticks.  In one
// tick every rollout outside its linear solve enqueues one phase of Newton
// work on its own stream, and all rollouts inside it take pcg_chunk
// iterations of one batched launch sequence on the server stream; then the
// host reads every rollout's scalars with one copy and moves each to its
// next phase, following minimize() exactly.

struct SlotRec {
    double scal[8];
    PcgCtl ctl;
    int flag, pad;
};
static_assert(S_COUNT <= 8, "slot scalars");

struct Request {
    long long id = 0;
    int N = 0, E = 0, ed = 0, material = 0, scheme = 0, method = 0;
    double grav[3] = {0, 0, 0}, damping = 0;
    std::vector<double> rest, youngs, poisson, density, bc_value, v0;
    std::vector<int> conn, bc_mask;
    double dt = 0;
    int steps = 0;
    using clock = std::chrono::steady_clock;
    clock::time_point t_submit, t_admit, t_resident;
    double est_bytes = 0, dev_bytes = 0;
    sors_ctx* ctx = nullptr;
    int status = SORS_OK;
    std::string err;
    std::vector<double> q;
    bool done = false;
    sors_request_stats stats{};
};

struct Member {
    Request* r = nullptr;
    sors_ctx* c = nullptr;
    int slot = -1;
    enum Phase { ASSEMBLE, PCG_START, PCG_RUN, POST, ENERGY } phase = ASSEMBLE;
    bool new_step = true, busy = false, brk = false;
    // aimed: the last solve's forcing term came from g_tol, so the next
    // residual check assembles the gradient alone (grad_only), as in minimize()
    bool aimed = false, grad_only = false;
    // options.reuse_tangent, as in minimize(): may_reuse for this Newton
    // iteration, reused when the last solve reused its tangent, check_failed
    // after a gradient-only check that neither converged nor reuses (the
    // tangent is assembled on the next tick)
    bool may_reuse = false, reused = false, check_failed = false;
    // the current solve runs in FP64: an all-FP64 server, or a mixed solve that
    // reached max_newton and goes on from its iterate in FP64 (as sors_step)
    bool fp64 = false;
    int step = 0, it = 0, bt = 0;
    long long launched = 0;
    double eta = 0, dxn = 0, gdx = 0, alpha = 1;
    sors_stats st{};
    PcgGrid g{};
};

double ctx_bytes(const sors_ctx* c) {
    const DevMem* all[] = {&c->conn,   &c->rest,  &c->mu,      &c->lam,   &c->mass,    &c->mask,    &c->bcval,
                           &c->q,      &c->qprev, &c->vprev,   &c->gprev, &c->gtmp,    &c->gint,    &c->g,
                           &c->fext,   &c->dxd,   &c->qtmp,    &c->rowptr, &c->cols,   &c->diag,    &c->rank,
                           &c->fullmap, &c->vals, &c->Minv,    &c->pr,    &c->pu,      &c->pp,      &c->ps,
                           &c->pw,     &c->px,    &c->partials, &c->scal, &c->pcgst,   &c->pcgctl,  &c->pcgpart,
                           &c->flag,   &c->sbase, &c->bcol,    &c->bcol16,   &c->selfslot, &c->csurf, &c->lagN,
                           &c->lagn,   &c->lagD,  &c->gcon,    &c->agroup, &c->afib,    &c->act,     &c->nu,
                           &c->incptr, &c->inc,  &c->teg,      &c->teb,    &c->tsptr,   &c->tsidx,   &c->tqslice, &c->vold,     &c->cdn,    &c->captr,   &c->canodes, &c->crot,
                           &c->cpairI, &c->cpairJ, &c->cpchunk, &c->cgptr, &c->cci,     &c->ccj,     &c->ccbase,
                           &c->cE,     &c->cEinv, &c->cEinv32, &c->crc,     &c->cX,      &c->cY,      &c->cpinv,   &c->cgpart,  &c->cpart,   &c->cinfo,
                           &c->adjb,   &c->adjq,  &c->adjv,    &c->adjE,   &c->adjA,   &c->mgroup, &c->mdir,
                           &c->mk,     &c->mact};
    double b = 0;
    for (const DevMem* d : all) b += (double)d->bytes;
    return b;
}

// device bytes of a rollout before its pattern exists (an upper estimate:
// node-block slices padded to 24 (tet) / 32 (hex) blocks per node)
double estimate_bytes(int N, int E, int ed, int coarse_nodes) {
    const double n = 3.0 * N, nblk = (double)N * (ed == 4 ? 24 : 32);
    // coarse correction: E and its inverse (nc^2 FP64 each) with the inverse's
    // FP32 copy (nc = 6 per aggregate, aggregates at most 2N / (nodes per
    // aggregate) and 512), the two Gauss-Jordan buffers (npad^2 FP64 each)
    // and per-node data
    double coarse = 0.0;
    if (coarse_nodes > 0) {
        const double nc = 6.0 * std::min(512.0, std::ceil(2.0 * N / coarse_target(N, coarse_nodes)));
        const double np = std::ceil(nc / kGjTile) * kGjTile;
        coarse = (2.0 * 8.0 + 4.0) * nc * nc + 2.0 * 8.0 * np * np + 24.0 * N;
    }
    // two-phase assembly: element rows (4 ed FP64 + ed^2 padded FP64 blocks at
    // worst) and the slot lists (ed^2 entries per element, one pointer per slot)
    const double rows = (double)E * (4 * ed * 8 + 10 * ed * ed * 8 + ed * ed * 4) + 4.0 * nblk;
    return n * (19 * 8 + 4) + 72.0 * N + nblk * (9 * 8 + 4 + 2 + 9 * 4) + (double)E * (ed * 4 + 16 + ed * ed) + rows +
           coarse + 8.0 * (5 * kMaxRedBlocks + 4 * kMaxRedBlocks) + (1 << 20);
}

}  // namespace

struct sors_server {
    sors_server_options opt{};
    int device = 0;
    std::string err;
    // shared state (mu)
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_prep;
    std::deque<Request*> pending;
    std::deque<Request*> ready;           // resident, not yet in a tick
    std::deque<std::function<void()>> tasks;    // prep / cleanup
    std::unordered_map<long long, std::unique_ptr<Request>> reqs;
    long long next_id = 1;
    int admitted = 0;                           // in prep, ready, active or cleanup
    double reserved = 0;
    bool stop = false, draining = false;
    sors_server_stats stats{};
    // worker-owned
    std::vector<Member> active;
    std::vector<int> free_slots;
    DevMem slots, bmem;
    void* h_stage = nullptr;                    // pinned: member table + launch tables
    SlotRec* h_slots = nullptr;           // pinned readback
    hipStream_t stream = nullptr;
    long long ticks = 0, launches = 0;         // published to stats under mu
    std::thread worker;
    std::vector<std::thread> prep;
};

namespace {

void srv_try_admit(sors_server* s) {  // mu held
    while (!s->pending.empty() && s->admitted < s->opt.max_active && !s->stop) {
        Request* r = s->pending.front();
        if (s->admitted > 0 && s->reserved + r->est_bytes > s->opt.hbm_budget_bytes) break;  // arrival order
        s->pending.pop_front();
        s->admitted++;
        s->reserved += r->est_bytes;
        r->t_admit = Request::clock::now();
        s->tasks.push_back([s, r] {
            sors_system_desc d{};
            d.n_nodes = r->N;
            d.n_elems = r->E;
            d.nodes_per_elem = r->ed;
            d.rest = r->rest.data();
            d.conn = r->conn.data();
            d.youngs = r->youngs.data();
            d.poisson = r->poisson.data();
            d.density = r->density.data();
            d.material = r->material;
            d.bc_mask = r->bc_mask.empty() ? nullptr : r->bc_mask.data();
            d.bc_value = r->bc_value.empty() ? nullptr : r->bc_value.data();
            for (int k = 0; k < 3; ++k) d.grav[k] = r->grav[k];
            d.damping_alpha = r->damping;
            d.scheme = r->scheme;
            d.method = r->method;
            sors_ctx* c = nullptr;
            int rc = sors_create(&d, &c);
            std::string msg = rc ? sors_last_error(nullptr) : "";
            if (rc == SORS_OK) {
                c->opt = s->opt.solver;
                try {
                    if (!r->v0.empty()) upload(c->vprev, r->v0.data(), (size_t)c->n, c->stream);
                    ensure_pattern(c, r->dt);
                    // allocated here so that the measured size below includes them
                    ensure_elem_rows(c, c->opt.precision == SORS_MIXED ? sizeof(float) : sizeof(double));
                    ensure_coarse(c);
                    HIPX(hipEventCreateWithFlags(&c->evPre, hipEventDisableTiming));
                    sync(c);
                } catch (const HipError& e) {
                    rc = SORS_ERR_HIP;
                    msg = e.msg;
                }
            }
            std::lock_guard<std::mutex> lk(s->mu);
            r->t_resident = Request::clock::now();
            if (rc != SORS_OK) {
                if (c) sors_destroy(c);
                r->status = rc;
                r->err = msg;
                s->reserved -= r->est_bytes;
                s->admitted--;
                s->stats.failed++;
                r->done = true;
                s->cv_done.notify_all();
                s->cv_work.notify_all();
                s->cv_prep.notify_all();
                srv_try_admit(s);
                return;
            }
            r->ctx = c;
            r->dev_bytes = ctx_bytes(c);
            s->reserved += r->dev_bytes - r->est_bytes;  // actual size from here on
            s->stats.resident_bytes = s->reserved;
            s->stats.peak_resident_bytes = std::max(s->stats.peak_resident_bytes, s->reserved);
            s->ready.push_back(r);
            s->cv_work.notify_all();
        });
        s->cv_prep.notify_one();
    }
}

// hand a finished (or failed) rollout to a prep thread: read q, free it
void srv_retire(sors_server* s, Member& m, int status, const std::string& msg) {
    Request* r = m.r;
    sors_ctx* c = m.c;
    c->scal_ext = nullptr;
    c->flag_ext = nullptr;
    s->free_slots.push_back(m.slot);
    r->stats.steps = m.step;
    r->stats.newton_iters = m.st.newton_iters;
    r->stats.pcg_iters = m.st.pcg_iters;
    r->stats.backtracks = m.st.backtracks;
    r->stats.fp64_fallbacks = m.st.fp64_fallbacks;
    std::lock_guard<std::mutex> lk(s->mu);
    s->tasks.push_front([s, r, c, status, msg] {
        int rc = status;
        std::string e = msg;
        if (rc == SORS_OK) {
            r->q.resize((size_t)c->n);
            if (sors_get_positions(c, r->q.data()) != SORS_OK) {
                rc = SORS_ERR_HIP;
                e = c->err;
            }
        }
        sors_destroy(c);
        std::lock_guard<std::mutex> lk2(s->mu);
        r->ctx = nullptr;
        r->status = rc;
        r->err = e;
        const auto now = Request::clock::now();
        r->stats.status = rc;
        r->stats.latency_ms = std::chrono::duration<double, std::milli>(now - r->t_submit).count();
        r->stats.queue_ms = std::chrono::duration<double, std::milli>(r->t_admit - r->t_submit).count();
        r->stats.prep_ms = std::chrono::duration<double, std::milli>(r->t_resident - r->t_admit).count();
        r->stats.device_bytes = r->dev_bytes;
        s->reserved -= r->dev_bytes;
        s->stats.resident_bytes = s->reserved;
        s->admitted--;
        if (rc == SORS_OK)
            s->stats.completed++;
        else
            s->stats.failed++;
        r->done = true;
        s->cv_done.notify_all();
        s->cv_work.notify_all();
        s->cv_prep.notify_all();
        srv_try_admit(s);
    });
    s->cv_prep.notify_one();
}

void srv_prep_loop(sors_server* s) {
    (void)hipSetDevice(s->device);
    for (;;) {
        std::function<void()> t;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_prep.wait(lk, [&] { return !s->tasks.empty() || (s->stop && s->admitted == 0); });
            if (s->tasks.empty()) return;
            t = std::move(s->tasks.front());
            s->tasks.pop_front();
        }
        t();
    }
}

// launch tables staged per tick: member indices and block prefixes of the
// starters (init, SpMV, update, coarse) and the runners (SpMV, update,
// coarse), M + 1 ints each, then the rollout counts of the four launch kinds
constexpr int kStageTables = 11;
enum { CNT_S = 0, CNT_SC, CNT_R, CNT_RC, kStageCounts };

// bytes staged per tick at precision HT, and the bytes of one precision's
// staging area (a multiple of 16, so the second area stays aligned)
template <typename HT>
size_t stage_bytes(int M) {
    return sizeof(BMember<HT>) * M + sizeof(int) * ((size_t)kStageTables * (M + 1) + kStageCounts);
}
size_t stage_area(int M) {
    static_assert(sizeof(BMember<double>) >= sizeof(BMember<float>), "FP64 member table is the larger");
    return (stage_bytes<double>(M) + 15) / 16 * 16;
}

// the batched PCG chunk of the rollouts in their linear solve at precision HT
// (starters begin their solve this tick)
template <typename HT>
void srv_batched_pcg(sors_server* s, const std::vector<int>& starters, const std::vector<int>& runners) {
    const int chunk = std::max(2, (s->opt.pcg_chunk + 1) / 2 * 2);
    {
        const int M = s->opt.max_active;
        // one staging area per precision: both may be in flight in one tick
        const size_t area = stage_area(M);
        const int which = sizeof(HT) == 8 ? 1 : 0;
        BMember<HT>* hm = (BMember<HT>*)((char*)s->h_stage + which * area);
        int* ht = (int*)(hm + M);
        int* tb[kStageTables];
        for (int k = 0; k < kStageTables; ++k) tb[k] = ht + k * (M + 1);
        int *idxS = tb[0], *tIS = tb[1], *tAS = tb[2], *tUS = tb[3], *idxCS = tb[4], *tCS = tb[5];
        int *idxR = tb[6], *tAR = tb[7], *tUR = tb[8], *idxCR = tb[9], *tCR = tb[10];
        for (int i : starters) {
            Member& m = s->active[i];
            BMember<HT>& b = hm[m.slot];
            b.v = pcg_view<HT>(m.c, m.eta);
            b.v.ctl = &((SlotRec*)s->slots.p)[m.slot].ctl;
            b.rhs = m.c->g.as<double>();
            b.part = m.c->pcgpart.as<double>();
            b.nbI = m.g.gn;
            b.nbA = m.g.gs;
            b.nbU = m.g.gu;
            // the rollout's coarse correction as its own pcg() would apply it
            b.nbC = m.c->coarse_ready ? m.c->nagg : 0;
            if (b.nbC) {
                b.ca = coarse_args(m.c);
                b.Einv = coarse_inverse<HT>(m.c);
                b.cpart = m.c->cpart.as<double>();
            }
        }
        int nsc = 0, nrc = 0;
        tIS[0] = tAS[0] = tUS[0] = tCS[0] = 0;
        for (size_t j = 0; j < starters.size(); ++j) {
            const Member& m = s->active[starters[j]];
            idxS[j] = m.slot;
            tIS[j + 1] = tIS[j] + m.g.gn;
            tAS[j + 1] = tAS[j] + m.g.gs;
            tUS[j + 1] = tUS[j] + m.g.gu;
            if (const int na = hm[m.slot].nbC) {
                idxCS[nsc] = m.slot;
                tCS[nsc + 1] = tCS[nsc] + na;
                ++nsc;
            }
        }
        tAR[0] = tUR[0] = tCR[0] = 0;
        for (size_t j = 0; j < runners.size(); ++j) {
            const Member& m = s->active[runners[j]];
            idxR[j] = m.slot;
            tAR[j + 1] = tAR[j] + m.g.gs;
            tUR[j + 1] = tUR[j] + m.g.gu;
            if (const int na = hm[m.slot].nbC) {
                idxCR[nrc] = m.slot;
                tCR[nrc + 1] = tCR[nrc] + na;
                ++nrc;
            }
        }
        int* cnt = ht + kStageTables * (M + 1);
        cnt[CNT_S] = (int)starters.size();
        cnt[CNT_SC] = nsc;
        cnt[CNT_R] = (int)runners.size();
        cnt[CNT_RC] = nrc;
        const size_t stage = stage_bytes<HT>(M);
        char* dbase = (char*)s->bmem.p + which * area;
        HIPX(hipMemcpyAsync(dbase, hm, stage, hipMemcpyHostToDevice, s->stream));
        const BMember<HT>* dm = (const BMember<HT>*)dbase;
        const int* dt_ = (const int*)(dbase + sizeof(BMember<HT>) * M);
        const int* db[kStageTables];
        for (int k = 0; k < kStageTables; ++k) db[k] = dt_ + k * (M + 1);
        const int *dIdxS = db[0], *dIS = db[1], *dAS = db[2], *dUS = db[3], *dIdxCS = db[4], *dCS = db[5];
        const int *dIdxR = db[6], *dAR = db[7], *dUR = db[8], *dIdxCR = db[9], *dCR = db[10];
        const int* dCnt = dt_ + kStageTables * (M + 1);
        const int ns = (int)starters.size(), nr = (int)runners.size();
        hipStream_t st = s->stream;
        // one workgroup per virtual block up to the persistent cap; beyond
        // it the workgroups walk the remaining virtual blocks
        auto grid = [](int total, int cap) { return std::min(total, cap); };
        // u += Z E^-1 Z^T r after an update (or the init) of the rollouts with a coarse correction
        auto coarse = [&](const int* didx, const int* dtab, const int* dcnt, int total, int par) {
            if (!total) return;
            const int gc = grid(total, kPersistCoarseBlocks);
            k_b_coarse_restrict<HT><<<gc, kRestrictThreads, 0, st>>>(dm, didx, dtab, dcnt);
            k_b_coarse_apply<HT><<<gc, kApplyThreads, 0, st>>>(dm, didx, dtab, dcnt, par);
            s->launches += 2;
        };
        if (ns) {
            for (int i : starters) HIPX(hipStreamWaitEvent(st, s->active[i].c->evPre, 0));
            const int gi = grid(tIS[ns], kPersistBlocks), ga = grid(tAS[ns], kPersistBlocks);
            const int gu = grid(tUS[ns], kPersistUpdBlocks);
            k_b_pcg_init<HT><<<gi, kBlock, 0, st>>>(dm, dIdxS, dIS, dCnt + CNT_S);
            coarse(dIdxCS, dCS, dCnt + CNT_SC, tCS[nsc], 0);
            k_b_pcg_spmv<HT><<<ga, kBlock, 0, st>>>(dm, dIdxS, dAS, dCnt + CNT_S, 0, 1);     // A_0
            k_b_pcg_update<HT><<<gu, kUpdBlock, 0, st>>>(dm, dIdxS, dUS, dCnt + CNT_S, 1, 1);  // U_1
            coarse(dIdxCS, dCS, dCnt + CNT_SC, tCS[nsc], 1);
            k_b_pcg_spmv<HT><<<ga, kBlock, 0, st>>>(dm, dIdxS, dAS, dCnt + CNT_S, 1, 0);     // A_1
            s->launches += 4;
        }
        // iterations k >= 2 of every rollout in its solve: every rollout has
        // completed an odd number of iterations, so parity follows j
        const int ga = grid(tAR[nr], kPersistBlocks), gu = grid(tUR[nr], kPersistUpdBlocks);
        for (int j = 0; j < chunk; ++j) {
            const int par = j & 1;
            k_b_pcg_update<HT><<<gu, kUpdBlock, 0, st>>>(dm, dIdxR, dUR, dCnt + CNT_R, par, 0);
            coarse(dIdxCR, dCR, dCnt + CNT_RC, tCR[nrc], par);
            k_b_pcg_spmv<HT><<<ga, kBlock, 0, st>>>(dm, dIdxR, dAR, dCnt + CNT_R, par, 0);
        }
        s->launches += 2 * chunk;
        HIPX(hipGetLastError());
        for (int i : starters) {
            s->active[i].phase = Member::PCG_RUN;
            s->active[i].launched = 1;
        }
        for (int i : runners) s->active[i].launched += chunk;
    }
}

// one tick over the resident rollouts; each rollout's kernels run in its own
// precision (member fp64), and the batched PCG once per precision in use
void srv_tick(sors_server* s) {
    const sors_options& o = s->opt.solver;
    const bool mixed = o.precision == SORS_MIXED;
    auto in_fp64 = [&](const Member& m) { return !mixed || m.fp64; };
    std::vector<int> starters[2], runners[2];  // [0] FP32 solves, [1] FP64 solves
    // 1. per-rollout phase work on the rollouts' streams
    for (int i = 0; i < (int)s->active.size(); ++i) {
        Member& m = s->active[i];
        sors_ctx* c = m.c;
        const double dt = m.r->dt;
        double* q = c->q.as<double>();
        m.busy = false;
        switch (m.phase) {
            case Member::ASSEMBLE:
                if (m.new_step) {
                    launch_predict(c, dt);
                    if (mixed) {
                        // the step's initial guess, for an FP64 restart after an inversion
                        ensure(c->qtmp, sizeof(double) * c->n);
                        HIPX(hipMemcpyAsync(c->qtmp.p, c->q.p, sizeof(double) * c->n, hipMemcpyDeviceToDevice,
                                            c->stream));
                    }
                    m.new_step = false;
                }
                clear_flag(c);
                m.may_reuse = !m.check_failed && may_reuse_tangent(o, m.it, m.reused, m.eta, m.st.gnorm);
                m.grad_only = !m.check_failed && (m.aimed || m.may_reuse);
                m.check_failed = false;
                if (in_fp64(m))
                    assemble<double>(c, q, dt, !m.grad_only, o.project_spd);
                else
                    assemble<float>(c, q, dt, !m.grad_only, o.project_spd);
                m.busy = true;
                break;
            case Member::POST:
                if (in_fp64(m))
                    enqueue_step_dots<double>(c, c->g.as<double>(), c->dxd.as<double>());
                else
                    enqueue_step_dots<float>(c, c->g.as<double>(), c->dxd.as<double>());
                if (c->method == SORS_MINIMIZE_NEWTON_LINESEARCH) enqueue_energy(c, q, c->dxd.as<double>(), 1.0, dt);
                m.busy = true;
                break;
            case Member::ENERGY:
                enqueue_energy(c, q, c->dxd.as<double>(), m.alpha, dt);
                m.busy = true;
                break;
            case Member::PCG_START:
                starters[in_fp64(m)].push_back(i);
                runners[in_fp64(m)].push_back(i);
                break;
            case Member::PCG_RUN:
                runners[in_fp64(m)].push_back(i);
                break;
        }
    }
    // 2. batched PCG on the server stream, FP32 solves then FP64 solves
    if (!runners[0].empty()) srv_batched_pcg<float>(s, starters[0], runners[0]);
    if (!runners[1].empty()) srv_batched_pcg<double>(s, starters[1], runners[1]);
    // 3. one readback of every slot
    for (Member& m : s->active)
        if (m.busy) HIPX(hipStreamSynchronize(m.c->stream));
    HIPX(hipMemcpyAsync(s->h_slots, s->slots.p, sizeof(SlotRec) * s->opt.max_active, hipMemcpyDeviceToHost,
                        s->stream));
    HIPX(hipStreamSynchronize(s->stream));
    s->ticks++;
    // 4. advance every rollout (minimize() / do_steps() restated as phases)
    std::vector<Member> keep;
    keep.reserve(s->active.size());
    for (Member& m : s->active) {
        sors_ctx* c = m.c;
        const SlotRec& sr = s->h_slots[m.slot];
        const double dt = m.r->dt;
        double* q = c->q.as<double>();
        int status = 1;  // 1: continue, 0: finished, < 0: error
        std::string msg;
        auto end_step = [&] {
            launch_update_state(c, dt);
            m.step++;
            if (m.step >= m.r->steps) {
                status = 0;
                return;
            }
            m.new_step = true;
            m.aimed = false;
            m.fp64 = false;
            m.it = 0;
            m.reused = m.check_failed = false;
            m.phase = Member::ASSEMBLE;
        };
        auto next_newton = [&] {
            if (++m.it >= o.max_newton) {
                if (mixed && !m.fp64) {
                    // as sors_step: the FP32 tangents did not bring this solve
                    // to the tolerances; it goes on from its iterate in FP64
                    m.fp64 = true;
                    m.st.fp64_fallbacks++;
                    m.it = 0;
                    m.reused = m.check_failed = false;
                    m.aimed = false;
                    m.phase = Member::ASSEMBLE;
                    return;
                }
                status = SORS_ERR_NOCONV;
                msg = "Newton's method did not converge";
                return;
            }
            m.phase = Member::ASSEMBLE;
        };
        auto accept = [&](double alpha) {
            k_axpy<<<red_grid(c->n, kBlock), kBlock, 0, c->stream>>>(c->n, alpha, c->dxd.as<double>(), q);
            m.st.dxnorm = alpha * m.dxn;
            if (m.st.dxnorm < o.dx_tol)
                end_step();
            else
                next_newton();
        };
        auto try_alpha = [&](double dE) {
            if (dE <= 0.1 * m.alpha * m.gdx) {
                accept(m.alpha);
                return;
            }
            m.alpha *= 0.5;
            m.st.backtracks++;
            if (++m.bt <= o.max_backtrack) {
                m.phase = Member::ENERGY;
                return;
            }
            if (m.alpha * 2 * m.dxn < o.dx_tol) {
                accept(m.alpha * 2);
            } else {
                status = SORS_ERR_LINESEARCH;
                msg = "line search failed to decrease the energy";
            }
        };
        switch (m.phase) {
            case Member::ASSEMBLE: {
                const double gg = sr.scal[S_GG];
                if (sr.flag & 1) {
                    if (mixed && !m.fp64) {
                        // as sors_step: a mixed solve that inverted an element
                        // restarts from the step's initial guess in FP64
                        HIPX(hipMemcpyAsync(c->q.p, c->qtmp.p, sizeof(double) * c->n, hipMemcpyDeviceToDevice,
                                            c->stream));
                        m.fp64 = true;
                        m.st.fp64_fallbacks++;
                        m.it = 0;
                        m.reused = m.check_failed = false;
                        m.aimed = false;
                        m.phase = Member::ASSEMBLE;
                        break;
                    }
                    status = SORS_ERR_INVERTED;
                    msg = "Inverted Elements";
                    break;
                }
                m.st.gnorm = std::sqrt(gg);
                if (!(gg == gg)) {
                    status = SORS_ERR_NAN;
                    msg = "Newton's method diverged (NaN residual)";
                    break;
                }
                if (o.g_tol > 0 && m.st.gnorm < o.g_tol) {
                    end_step();
                    break;
                }
                const bool reuse = m.grad_only && m.may_reuse && in_reach(o, m.it, m.st.gnorm);
                if (m.grad_only && !reuse) {  // the gradient-only check failed: the tangent on the next tick
                    m.aimed = false;
                    m.check_failed = true;
                    break;
                }
                const double base = forcing_base(o, m.it);
                double eta = base;
                if (o.g_tol > 0 && m.st.gnorm > 0) eta = std::min(0.1, std::max(eta, kAim * o.g_tol / m.st.gnorm));
                m.eta = eta;
                m.aimed = o.g_tol > 0 && eta > base;
                m.reused = reuse;
                // preconditioner only for a tangent that is solved (as minimize()
                // does; a reused tangent keeps its own); the next tick's batched
                // launches wait for it
                if (!reuse) {
                    if (in_fp64(m))
                        build_preconditioner<double>(c);
                    else
                        build_preconditioner<float>(c);
                }
                HIPX(hipEventRecord(c->evPre, c->stream));
                m.phase = Member::PCG_START;
                break;
            }
            case Member::PCG_RUN:
                if (sr.ctl.done || m.launched >= o.pcg_max_iter) {
                    m.st.pcg_iters += sr.ctl.done ? sr.ctl.iter : m.launched;
                    m.brk = sr.ctl.done && sr.ctl.breakdown;
                    m.phase = Member::POST;
                }
                break;
            case Member::POST: {
                m.st.newton_iters++;
                m.dxn = std::sqrt(sr.scal[S_DD]);
                m.gdx = sr.scal[S_GD];
                if (!(m.dxn == m.dxn)) {
                    status = SORS_ERR_NAN;
                    msg = "Newton's method diverged";
                    break;
                }
                if (m.brk && m.dxn == 0.0) {
                    status = SORS_ERR_NAN;
                    msg = "PCG breakdown (indefinite tangent)";
                    break;
                }
                if (c->method != SORS_MINIMIZE_NEWTON_LINESEARCH || m.dxn < o.dx_tol) {
                    k_axpy<<<red_grid(c->n, kBlock), kBlock, 0, c->stream>>>(c->n, 1.0, c->dxd.as<double>(), q);
                    m.st.dxnorm = m.dxn;
                    if (m.dxn < o.dx_tol)
                        end_step();
                    else
                        next_newton();
                    break;
                }
                m.alpha = 1.0;
                m.bt = 0;
                try_alpha(energy_sum(c, sr.scal + S_EELEM));
                break;
            }
            case Member::ENERGY:
                try_alpha(energy_sum(c, sr.scal + S_EELEM));
                break;
            default:
                break;
        }
        if (status == 1)
            keep.push_back(m);
        else
            srv_retire(s, m, status == 0 ? SORS_OK : status, msg);
    }
    s->active.swap(keep);
}

void srv_worker(sors_server* s) {
    (void)hipSetDevice(s->device);
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->active.empty())
                s->cv_work.wait(lk, [&] { return !s->ready.empty() || (s->stop && s->admitted == 0); });
            if (s->active.empty() && s->ready.empty() && s->stop && s->admitted == 0) return;
            while (!s->ready.empty() && !s->free_slots.empty()) {
                Request* r = s->ready.front();
                s->ready.pop_front();
                Member m;
                m.r = r;
                m.c = r->ctx;
                m.slot = s->free_slots.back();
                s->free_slots.pop_back();
                SlotRec* sl = (SlotRec*)s->slots.p + m.slot;
                m.c->scal_ext = sl->scal;
                m.c->flag_ext = &sl->flag;
                m.g = pcg_grid(m.c);
                s->active.push_back(m);
            }
            s->stats.ticks = s->ticks;
            s->stats.batched_launches = s->launches;
            s->stats.active = (int)s->active.size();
            s->stats.peak_active = std::max(s->stats.peak_active, s->stats.active);
        }
        if (s->active.empty()) continue;
        try {
            srv_tick(s);
        } catch (const HipError& e) {
            std::vector<Member> all;
            all.swap(s->active);
            for (Member& m : all) srv_retire(s, m, SORS_ERR_HIP, e.msg);
            std::lock_guard<std::mutex> lk(s->mu);
            s->err = e.msg;
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

const char* sors_last_error(const sors_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

void sors_default_options(sors_options* o) {
    o->precision = SORS_MIXED;
    o->project_spd = 1;
    o->max_newton = 100;
    o->dx_tol = 1e-9;
    o->g_tol = 0.0;
    o->pcg_rtol = 1e-6;
    o->pcg_max_iter = 20000;
    o->max_backtrack = 60;
    o->pcg_check_every = 32;
    o->use_graph = 1;
    o->profile_kernels = 0;
    o->predict = 2;
    o->coarse_nodes = 2048;
    o->coarse_every = 16;
    o->pcg_rtol_first = 0.0;
    o->reuse_tangent = 0;
}

int sors_create(const sors_system_desc* d, sors_ctx** out) {
    g_create_error.clear();
    if (!d || !out) {
        g_create_error = "null argument";
        return SORS_ERR_ARG;
    }
    *out = nullptr;
    if (d->nodes_per_elem != 4 && d->nodes_per_elem != 8) {
        g_create_error = "nodes_per_elem must be 4 (Tetrahedron) or 8 (Hexahedron)";
        return SORS_ERR_ARG;
    }
    if (d->material != SORS_NEOHOOKEAN && d->material != SORS_TEST_POLY && d->material != SORS_STABLE_NEOHOOKEAN) {
        g_create_error = "Unknown energy type (material " + std::to_string(d->material) + ")";
        return SORS_ERR_ARG;
    }
    if (d->material == SORS_TEST_POLY && d->nodes_per_elem != 4) {
        g_create_error = "the polynomial test energy is defined for tetrahedra";
        return SORS_ERR_ARG;
    }
    if (d->n_nodes <= 0 || d->n_elems < 0 || !d->rest || (d->n_elems > 0 && !d->conn)) {
        g_create_error = "empty or missing mesh arrays";
        return SORS_ERR_ARG;
    }
    if (d->scheme != SORS_BACKWARD_EULER && d->scheme != SORS_CRANK_NICOLSON) {
        g_create_error = "Unknown time stepping scheme";
        return SORS_ERR_ARG;
    }
    if (d->method < SORS_MINIMIZE_NEWTON || d->method > SORS_MINIMIZE_FROZENHESSIAN) {
        g_create_error = "Solver method not recognized";
        return SORS_ERR_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        g_create_error = "no HIP device available";
        return SORS_ERR_HIP;
    }
    std::unique_ptr<sors_ctx> c(new sors_ctx());
    try {
        c->N = d->n_nodes;
        c->E = d->n_elems;
        c->ed = d->nodes_per_elem;
        c->n = 3 * c->N;
        c->material = d->material;
        c->scheme = d->scheme;
        c->method = d->method;
        for (int k = 0; k < 3; ++k) c->grav[k] = d->grav[k];
        c->damping = d->damping_alpha;
        sors_default_options(&c->opt);
        HIPX(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        {
            // the allocation-free entry point's scratch: its address on the
            // current device, looked up here (once per device) with that
            // device current; a failure is reported by that entry point
            int dev = -1;
            HIPX(hipGetDevice(&dev));
            c->scratch_dev = scratch_address(dev, &c->scratch_err);
        }
        const int N = c->N, E = c->E, ed = c->ed, n = c->n;
        c->h_conn.assign(d->conn, d->conn + (size_t)E * ed);
        c->h_rest.assign(d->rest, d->rest + (size_t)n);
        for (size_t i = 0; i < c->h_conn.size(); ++i)
            if (c->h_conn[i] < 0 || c->h_conn[i] >= N) {
                g_create_error = "element index out of range";
                return SORS_ERR_ARG;
            }
        c->has_mask = false;
        c->h_mask.assign(n, 0);
        if (d->bc_mask)
            for (int i = 0; i < n; ++i) {
                c->h_mask[i] = d->bc_mask[i] != 0;
                c->has_mask |= c->h_mask[i] != 0;
            }
        // element SoA connectivity, Lame constants (neoHookean.cpp:10-11),
        // lumped mass rho V0 / ed per DOF (energy.cpp:176-189)
        std::vector<int> connT((size_t)E * ed);
        std::vector<double> mu(E), lam(E), mass(n, 0.0);
        c->h_nu.assign(E, 0.0);
        if (d->poisson) std::copy(d->poisson, d->poisson + E, c->h_nu.begin());
        for (int e = 0; e < E; ++e) {
            for (int k = 0; k < ed; ++k) connT[(size_t)k * E + e] = c->h_conn[(size_t)e * ed + k];
            const double Ey = d->youngs ? d->youngs[e] : 0.0, nu = d->poisson ? d->poisson[e] : 0.0;
            material_constants(c->material, Ey, nu, &mu[e], &lam[e]);
            const int* cn = &c->h_conn[(size_t)e * ed];
            auto P = [&](int k) { return d->rest + 3 * (size_t)cn[k]; };
            auto tv = [&](int a, int b, int cc, int dd) {
                const double *A = P(a), *B = P(b), *C = P(cc), *D = P(dd);
                double ab[3], ac[3], ad[3];
                for (int t = 0; t < 3; ++t) {
                    ab[t] = B[t] - A[t];
                    ac[t] = C[t] - A[t];
                    ad[t] = D[t] - A[t];
                }
                return std::fabs(ab[0] * (ac[1] * ad[2] - ac[2] * ad[1]) + ab[1] * (ac[2] * ad[0] - ac[0] * ad[2]) +
                                 ab[2] * (ac[0] * ad[1] - ac[1] * ad[0])) /
                       6.0;
            };
            double V0;
            if (ed == 4) {
                V0 = tv(0, 1, 2, 3);
            } else {
                static const int T5[5][4] = {{0, 1, 3, 4}, {1, 2, 3, 6}, {1, 4, 5, 6}, {3, 4, 6, 7}, {1, 3, 4, 6}};
                V0 = 0;
                for (int t = 0; t < 5; ++t) V0 += tv(T5[t][0], T5[t][1], T5[t][2], T5[t][3]);
            }
            if (!(V0 > 0.0) && c->material != SORS_TEST_POLY) {
                g_create_error = "Volume of element must be positive (element " + std::to_string(e) + ")";
                return SORS_ERR_ARG;
            }
            const double me = (d->density ? d->density[e] : 0.0) * V0;
            for (int k = 0; k < ed; ++k)
                for (int t = 0; t < 3; ++t) mass[3 * (size_t)cn[k] + t] += me / ed;
        }
        hipStream_t st = c->stream;
        upload(c->conn, connT.data(), connT.size(), st);
        upload(c->rest, d->rest, (size_t)n, st);
        upload(c->mu, mu.data(), mu.size(), st);
        upload(c->lam, lam.data(), lam.size(), st);
        upload(c->mass, mass.data(), mass.size(), st);
        upload(c->mask, c->h_mask.data(), c->h_mask.size(), st);
        std::vector<double> bcv(n, 0.0);
        if (d->bc_value) std::copy(d->bc_value, d->bc_value + n, bcv.begin());
        upload(c->bcval, bcv.data(), bcv.size(), st);
        upload(c->qprev, d->rest, (size_t)n, st);
        upload(c->q, d->rest, (size_t)n, st);
        std::vector<double> zeros(n, 0.0);
        upload(c->vprev, zeros.data(), n, st);
        upload(c->gprev, zeros.data(), n, st);
        upload(c->gtmp, zeros.data(), n, st);
        ensure(c->gint, sizeof(double) * n);
        ensure(c->g, sizeof(double) * n);
        ensure(c->dxd, sizeof(double) * n);
        ensure(c->qtmp, sizeof(double) * n);
        for (DevMem* v : {&c->pr, &c->pu, &c->pp, &c->ps, &c->pw, &c->px}) ensure(*v, sizeof(double) * n);
        ensure(c->Minv, sizeof(double) * 9 * (size_t)N);
        ensure(c->partials, sizeof(double) * kMaxRedBlocks * 4);
        ensure(c->pcgpart, sizeof(double) * 5 * kMaxRedBlocks);
        ensure(c->pcgctl, sizeof(PcgCtl));
        ensure(c->scal, sizeof(double) * S_COUNT);
        ensure(c->pcgst, 2 * sizeof(PcgState));
        ensure(c->flag, sizeof(int));
        fill<double>(c->S(), S_COUNT, 0.0, st);
        fill<int>(c->F(), 1, 0, st);
        HIPX(hipHostMalloc((void**)&c->h_pin, 4096, hipHostMallocDefault));
        HIPX(hipStreamSynchronize(st));
    } catch (const HipError& e) {
        g_create_error = e.msg;
        sors_destroy(c.release());
        return SORS_ERR_HIP;
    }
    *out = c.release();
    return SORS_OK;
}

void sors_destroy(sors_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->h_pin) (void)hipHostFree(c->h_pin);
    for (hipEvent_t e : c->pev) (void)hipEventDestroy(e);
    for (auto& row : c->gexec)
        for (hipGraphExec_t g : row)
            if (g) (void)hipGraphExecDestroy(g);
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->evG) (void)hipEventDestroy(c->evG);
    if (c->evPre) (void)hipEventDestroy(c->evPre);
    if (c->evR) (void)hipEventDestroy(c->evR);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    c->stream = nullptr;
    delete c;
}

int sors_set_options(sors_ctx* c, const sors_options* o) {
    if (!c || !o) return SORS_ERR_ARG;
    if (o->precision != SORS_FP64 && o->precision != SORS_MIXED) return fail(c, SORS_ERR_ARG, "bad precision");
    if (o->max_newton < 1 || o->pcg_max_iter < 1 || o->max_backtrack < 0)
        return fail(c, SORS_ERR_ARG, "iteration limits must be positive");
    if (o->coarse_nodes < 0) return fail(c, SORS_ERR_ARG, "coarse_nodes must be >= 0");
    if (!(o->pcg_rtol_first >= 0.0 && o->pcg_rtol_first < 1.0))
        return fail(c, SORS_ERR_ARG, "pcg_rtol_first must be in [0, 1)");
    if (o->reuse_tangent != 0 && o->reuse_tangent != 1) return fail(c, SORS_ERR_ARG, "reuse_tangent must be 0 or 1");
    c->opt = *o;
    return SORS_OK;
}

int sors_get_options(const sors_ctx* c, sors_options* o) {
    if (!c || !o) return SORS_ERR_ARG;
    *o = c->opt;
    return SORS_OK;
}

#define SORS_GUARD(c)                                                   \
    if (!(c)) return SORS_ERR_ARG;                                      \
    try {
#define SORS_GUARD_END(c)                                               \
    }                                                                   \
    catch (const HipError& e) {                                         \
        return fail(c, SORS_ERR_HIP, e.msg);                            \
    }

int sors_set_initial_velocity(sors_ctx* c, const double* v) {
    SORS_GUARD(c)
    c->vold_valid = false;
    if (!v) return fail(c, SORS_ERR_ARG, "null velocity");
    upload(c->vprev, v, (size_t)c->n, c->stream);
    sync(c);
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_set_initial_deformation(sors_ctx* c, const double* q) {
    SORS_GUARD(c)
    c->vold_valid = false;
    if (!q) return fail(c, SORS_ERR_ARG, "null positions");
    upload(c->qprev, q, (size_t)c->n, c->stream);
    upload(c->q, q, (size_t)c->n, c->stream);
    sync(c);
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_set_external_force(sors_ctx* c, const double* f) {
    SORS_GUARD(c)
    if (!f) {
        c->has_fext = false;
        return SORS_OK;
    }
    upload(c->fext, f, (size_t)c->n, c->stream);
    sync(c);
    c->has_fext = true;
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_get_state(sors_ctx* c, double* qprev, double* vprev) {
    SORS_GUARD(c)
    if (qprev) HIPX(hipMemcpyAsync(qprev, c->qprev.p, sizeof(double) * c->n, hipMemcpyDeviceToHost, c->stream));
    if (vprev) HIPX(hipMemcpyAsync(vprev, c->vprev.p, sizeof(double) * c->n, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_update_state(sors_ctx* c, const double* q, double dt) {
    SORS_GUARD(c)
    if (!q) return fail(c, SORS_ERR_ARG, "null positions");
    if (!(dt > 0.0)) return fail(c, SORS_ERR_ARG, "dt must be > 0");
    upload(c->q, q, (size_t)c->n, c->stream);
    launch_update_state(c, dt);
    HIPX(hipGetLastError());
    std::memcpy(c->cargs.dc_last, c->cargs.dc, sizeof(c->cargs.dc));
    sync(c);
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_minimize(sors_ctx* c, double* q, double dt, int method, sors_stats* st) {
    SORS_GUARD(c)
    if (!q) return fail(c, SORS_ERR_ARG, "null positions");
    if (!(dt > 0.0)) return fail(c, SORS_ERR_ARG, "dt must be > 0");
    if (method != SORS_MINIMIZE_NEWTON && method != SORS_MINIMIZE_NEWTON_LINESEARCH &&
        method != SORS_MINIMIZE_FROZENHESSIAN && method != SORS_MINIMIZE_SQP)
        return fail(c, SORS_ERR_ARG, "method must be minimize_newton, minimize_newton_linesearch, "
                                     "minimize_frozenhessian or minimize_SQP");
    sors_stats local{};
    if (!st) st = &local;
    *st = sors_stats{};
    struct MethodReset {
        sors_ctx* c;
        int m;
        ~MethodReset() { c->method = m; }
    } reset{c, c->method};
    c->method = method;
    auto t0 = std::chrono::steady_clock::now();
    upload(c->q, q, (size_t)c->n, c->stream);
    ensure_pattern(c, dt);
    if (!c->sqp()) contact_lag(c);
    const int rc = minimize_step(c, dt, st);
    if (rc != SORS_OK) return rc;
    HIPX(hipMemcpyAsync(q, c->q.p, sizeof(double) * c->n, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    st->ms_total = ms_since(t0);
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_get_coarse(sors_ctx* c, int* nc, double* E, double* Einv) {
    SORS_GUARD(c)
    if (!nc) return fail(c, SORS_ERR_ARG, "null nc");
    *nc = c->coarse_ready ? c->nc : 0;
    if (!c->coarse_ready) return SORS_OK;
    const size_t nc2 = (size_t)c->nc * c->nc;
    if (E) HIPX(hipMemcpyAsync(E, c->cE.p, sizeof(double) * nc2, hipMemcpyDeviceToHost, c->stream));
    if (Einv) HIPX(hipMemcpyAsync(Einv, c->cEinv.p, sizeof(double) * nc2, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_get_positions(sors_ctx* c, double* q) {
    SORS_GUARD(c)
    if (!q) return fail(c, SORS_ERR_ARG, "null output");
    HIPX(hipMemcpyAsync(q, c->q.p, sizeof(double) * c->n, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_step(sors_ctx* c, double* q, double dt, int substeps, sors_stats* st) {
    SORS_GUARD(c)
    if (!q) return fail(c, SORS_ERR_ARG, "null positions");
    upload(c->q, q, (size_t)c->n, c->stream);
    int rc = do_steps(c, dt, substeps, st);
    if (rc != SORS_OK) return rc;
    HIPX(hipMemcpyAsync(q, c->q.p, sizeof(double) * c->n, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_step_resident(sors_ctx* c, double dt, int substeps, sors_stats* st) {
    SORS_GUARD(c)
    return do_steps(c, dt, substeps, st);
    SORS_GUARD_END(c)
}

int sors_energy(sors_ctx* c, const double* q, double dt, double* out) {
    SORS_GUARD(c)
    if (!q || !out) return fail(c, SORS_ERR_ARG, "null argument");
    upload(c->qtmp, q, (size_t)c->n, c->stream);
    clear_flag(c);
    *out = energy_abs(c, c->qtmp.as<double>(), dt);
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_energy_noalloc(sors_ctx* c, const double* q_dev, double dt, double* energy_dev, void* stream) {
    SORS_GUARD(c)
    if (!q_dev || !energy_dev) return fail(c, SORS_ERR_ARG, "null argument");
    if (!c->scratch_dev) return fail(c, SORS_ERR_HIP, "static scratch unavailable: " + c->scratch_err);
    if (c->has_pressure)
        return fail(c, SORS_ERR_UNSUPPORTED, "pressure forces need the context's own buffers (use sors_energy)");
    // the three reductions run with red_grid() blocks each, capped at
    // kMaxRedBlocks; the scratch holds kScratchPartials partials
    const long long work = std::max<long long>(std::max<long long>(c->E, c->n), c->has_contact ? c->cargs.nsurf : 0);
    if (red_grid(work, kBlock) > kScratchPartials)
        return fail(c, SORS_ERR_ARG, "energy reduction needs " + std::to_string(red_grid(work, kBlock)) +
                                         " partial sums; the static scratch holds " +
                                         std::to_string(kScratchPartials));
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    double* sc = c->scratch_dev;
    int* flag = (int*)(sc + kScratchFlag);
    // the flag is only ever or-ed into: start it at zero (a fill kernel, not
    // a memset); partials and sums are overwritten by every call
    k_fill<int><<<1, 64, 0, st>>>(flag, 1, 0);
    enqueue_energy(c, q_dev, nullptr, 0.0, dt, EnergyTarget{st, sc, sc + kScratchSums, flag});
    k_energy_total<<<1, 64, 0, st>>>(sc + kScratchSums, penalty_on(c) ? 1 : 0, energy_dev);
    HIPX(hipGetLastError());
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_gradient(sors_ctx* c, const double* q, double dt, double* g) {
    SORS_GUARD(c)
    if (!q || !g) return fail(c, SORS_ERR_ARG, "null argument");
    upload(c->qtmp, q, (size_t)c->n, c->stream);
    ensure_pattern(c, dt);
    clear_flag(c);
    assemble<double>(c, c->qtmp.as<double>(), dt, false, 0);
    HIPX(hipMemcpyAsync(g, c->g.p, sizeof(double) * c->n, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_cfl_timestep(sors_ctx* c, const double* q, double dt, double cfl, double* dt_cfl) {
    SORS_GUARD(c)
    if (!dt_cfl) return fail(c, SORS_ERR_ARG, "null argument");
    if (c->ed != 4 && c->ed != 8) return fail(c, SORS_ERR_UNSUPPORTED, "CFL time step: tet4 and hex8 only");
    const double* qd = c->q.as<double>();
    if (q) {
        upload(c->qtmp, q, (size_t)c->n, c->stream);
        qd = c->qtmp.as<double>();
    }
    const int nb = red_grid(std::max<long long>(c->E, c->n), kBlock);
    double* part = c->partials.as<double>();
    if (c->ed == 4)
        k_cfl_partials<4><<<nb, kBlock, 0, c->stream>>>(c->E, c->conn.as<int>(), qd, c->n, c->vprev.as<double>(), part);
    else
        k_cfl_partials<8><<<nb, kBlock, 0, c->stream>>>(c->E, c->conn.as<int>(), qd, c->n, c->vprev.as<double>(), part);
    HIPX(hipGetLastError());
    std::vector<double> h((size_t)kMaxRedBlocks + nb);
    HIPX(hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    sync(c);
    double emin = INFINITY, vmax = 0.0;
    for (int b = 0; b < nb; ++b) {
        emin = std::fmin(emin, h[b]);
        vmax = std::fmax(vmax, h[(size_t)kMaxRedBlocks + b]);
    }
    // solver.cpp:325-330 (no motion or no elements: dt)
    const double dtc = cfl * emin / vmax;
    *dt_cfl = dtc < dt ? dtc : dt;
    return SORS_OK;
    SORS_GUARD_END(c)
}

long long sors_pattern_nnz(sors_ctx* c, double dt) {
    if (!c) return SORS_ERR_ARG;
    try {
        ensure_pattern(c, dt);
        return c->pat.nnz;
    } catch (const HipError& e) {
        return fail(c, SORS_ERR_HIP, e.msg);
    }
}

int sors_get_pattern(sors_ctx* c, double dt, int* rowptr, int* cols) {
    SORS_GUARD(c)
    ensure_pattern(c, dt);
    // read back what the kernels use
    if (rowptr) HIPX(hipMemcpyAsync(rowptr, c->rowptr.p, sizeof(int) * (c->n + 1), hipMemcpyDeviceToHost, c->stream));
    if (cols) HIPX(hipMemcpyAsync(cols, c->cols.p, sizeof(int) * c->pat.nnz, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_get_scatter_map(sors_ctx* c, double dt, int* map) {
    SORS_GUARD(c)
    if (!map) return fail(c, SORS_ERR_ARG, "null output");
    ensure_pattern(c, dt);
    // expand the device-resident compact map with the kernels' index rule
    HostPattern p;
    p.node_granular = c->pat.node_granular;
    p.rowptr.resize(c->n + 1);
    HIPX(hipMemcpyAsync(p.rowptr.data(), c->rowptr.p, sizeof(int) * (c->n + 1), hipMemcpyDeviceToHost, c->stream));
    if (p.node_granular) {
        p.rank.resize(c->pat.rank.size());
        HIPX(hipMemcpyAsync(p.rank.data(), c->rank.p, p.rank.size(), hipMemcpyDeviceToHost, c->stream));
        sync(c);
        scatter_map_full(p, c->E, c->ed, c->h_conn.data(), c->has_mask ? c->h_mask.data() : nullptr, c->sqp(), map);
    } else {
        HIPX(hipMemcpyAsync(map, c->fullmap.p, sizeof(int) * c->pat.fullmap.size(), hipMemcpyDeviceToHost,
                            c->stream));
        sync(c);
    }
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_hessian(sors_ctx* c, const double* q, double dt, int precision, int project, double* vals) {
    SORS_GUARD(c)
    if (!q || !vals) return fail(c, SORS_ERR_ARG, "null argument");
    upload(c->qtmp, q, (size_t)c->n, c->stream);
    ensure_pattern(c, dt);
    clear_flag(c);
    const long long nnz = c->pat.nnz;
    struct ExportReset {
        sors_ctx* c;
        ~ExportReset() { c->export_pressure_jacobian = false; }
    } export_reset{c};
    c->export_pressure_jacobian = true;
    if (precision == SORS_MIXED) {
        assemble<float>(c, c->qtmp.as<double>(), dt, true, project);
    } else {
        assemble<double>(c, c->qtmp.as<double>(), dt, true, project);
    }
    c->export_pressure_jacobian = false;
    sync(c);
    if (c->block()) {
        DevMem out;
        ensure(out, sizeof(double) * std::max<long long>(nnz, 1));
        const int* mk = c->has_mask && !c->sqp() ? c->mask.as<int>() : nullptr;
        if (precision == SORS_MIXED)
            k_block_to_csr<float><<<grid_for(c->n), kBlock, 0, c->stream>>>(
                c->n, c->rowptr.as<int>(), c->cols.as<int>(), c->sbase.as<int>(), c->selfslot.as<int>(), mk,
                c->vals.as<float>(), out.as<double>());
        else
            k_block_to_csr<double><<<grid_for(c->n), kBlock, 0, c->stream>>>(
                c->n, c->rowptr.as<int>(), c->cols.as<int>(), c->sbase.as<int>(), c->selfslot.as<int>(), mk,
                c->vals.as<double>(), out.as<double>());
        HIPX(hipGetLastError());
        HIPX(hipMemcpyAsync(vals, out.p, sizeof(double) * nnz, hipMemcpyDeviceToHost, c->stream));
        sync(c);
    } else if (precision == SORS_MIXED) {
        std::vector<float> tmp(nnz);
        HIPX(hipMemcpyAsync(tmp.data(), c->vals.p, sizeof(float) * nnz, hipMemcpyDeviceToHost, c->stream));
        sync(c);
        for (long long i = 0; i < nnz; ++i) vals[i] = tmp[i];
    } else {
        HIPX(hipMemcpyAsync(vals, c->vals.p, sizeof(double) * nnz, hipMemcpyDeviceToHost, c->stream));
        sync(c);
    }
    const int fl = read_flag(c);
    if (fl & 1) return fail(c, SORS_ERR_INVERTED, "Inverted Elements");
    if (fl & 4) return fail(c, SORS_ERR_ARG, "a pressure face joins vertices that share no element");
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_element_eval(sors_ctx* c, const double* q, int precision, int project, double* energy, double* grad,
                      double* hess) {
    SORS_GUARD(c)
    if (!q) return fail(c, SORS_ERR_ARG, "null positions");
    upload(c->qtmp, q, (size_t)c->n, c->stream);
    const int nd = 3 * c->ed;
    DevMem de, dg, dh;
    ensure(de, sizeof(double) * std::max(1, c->E));
    ensure(dg, sizeof(double) * std::max<size_t>(1, (size_t)c->E * nd));
    ensure(dh, sizeof(double) * std::max<size_t>(1, (size_t)c->E * nd * nd));
    ElemArgs a = elem_args(c, c->qtmp.as<double>());
    a.ex_energy = de.as<double>();
    a.ex_grad = dg.as<double>();
    a.ex_hess = dh.as<double>();
    clear_flag(c);
    if (precision == SORS_MIXED)
        launch_elem<float, 2>(c, a, project);
    else
        launch_elem<double, 2>(c, a, project);
    if (energy) HIPX(hipMemcpyAsync(energy, de.p, sizeof(double) * c->E, hipMemcpyDeviceToHost, c->stream));
    if (grad) HIPX(hipMemcpyAsync(grad, dg.p, sizeof(double) * (size_t)c->E * nd, hipMemcpyDeviceToHost, c->stream));
    if (hess)
        HIPX(hipMemcpyAsync(hess, dh.p, sizeof(double) * (size_t)c->E * nd * nd, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_linear_solve(sors_ctx* c, const double* q, double dt, const double* b, double* x, double rtol, int max_iter,
                      int* iters) {
    SORS_GUARD(c)
    if (!q || !b || !x) return fail(c, SORS_ERR_ARG, "null argument");
    upload(c->qtmp, q, (size_t)c->n, c->stream);
    upload(c->dxd, b, (size_t)c->n, c->stream);
    ensure_pattern(c, dt);
    clear_flag(c);
    bool brk = false;
    long long it;
    if (c->opt.precision == SORS_MIXED) {
        assemble<float>(c, c->qtmp.as<double>(), dt, true, c->opt.project_spd);
        build_preconditioner<float>(c);
        it = pcg<float>(c, c->dxd.as<double>(), 1.0, rtol, max_iter, &brk);
        k_to_f64<float><<<red_grid(c->n, kBlock), kBlock, 0, c->stream>>>(c->n, c->px.as<float>(), c->dxd.as<double>());
    } else {
        assemble<double>(c, c->qtmp.as<double>(), dt, true, c->opt.project_spd);
        build_preconditioner<double>(c);
        it = pcg<double>(c, c->dxd.as<double>(), 1.0, rtol, max_iter, &brk);
        k_to_f64<double><<<red_grid(c->n, kBlock), kBlock, 0, c->stream>>>(c->n, c->px.as<double>(), c->dxd.as<double>());
    }
    HIPX(hipMemcpyAsync(x, c->dxd.p, sizeof(double) * c->n, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    if (iters) *iters = (int)it;
    if (brk) return fail(c, SORS_ERR_NAN, "PCG breakdown (indefinite tangent)");
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_bench_kernel(sors_ctx* c, int kernel, int reps, double* avg_ms, double* algo_bytes) {
    SORS_GUARD(c)
    if (reps < 1 || !avg_ms) return fail(c, SORS_ERR_ARG, "reps must be >= 1");
    if (!c->pat_on_device) return fail(c, SORS_ERR_ARG, "no assembled system yet: run a step first");
    const bool mixed = c->opt.precision == SORS_MIXED;
    const double hb = mixed ? 4.0 : 8.0;
    const int N = c->N, n = c->n;
    const long long nnz = c->pat.nnz;
    // bench scalars: a solve that never converges and never breaks down
    ensure(c->pcgst_bench, 2 * sizeof(PcgState) + sizeof(PcgCtl));
    PcgState* bst = c->pcgst_bench.as<PcgState>();
    PcgCtl* bctl = (PcgCtl*)(bst + 2);
    const PcgGrid g = pcg_grid(c);
    PcgState init[2] = {};
    for (auto& t : init) {
        t.gamma = (double)g.gu;
        t.alpha = 1e30;
        t.tol2 = -1.0;
        t.rr0 = 1.0;
    }
    PcgCtl cinit{};
    hipEvent_t* ev = c->events(2);
    double bytes = 0;
    auto run = [&](auto launch) {
        fill<double>(c->pcgpart.as<double>(), 5LL * kMaxRedBlocks, 1.0, c->stream);
        HIPX(hipMemcpyAsync(bst, init, sizeof(init), hipMemcpyHostToDevice, c->stream));
        HIPX(hipMemcpyAsync(bctl, &cinit, sizeof(cinit), hipMemcpyHostToDevice, c->stream));
        launch();  // warm
        sync(c);
        HIPX(hipEventRecord(ev[0], c->stream));
        for (int i = 0; i < reps; ++i) launch();
        HIPX(hipEventRecord(ev[1], c->stream));
        HIPX(hipEventSynchronize(ev[1]));
        HIPX(hipGetLastError());
        float ms = 0;
        HIPX(hipEventElapsedTime(&ms, ev[0], ev[1]));
        *avg_ms = ms / reps;
    };
    if (kernel == SORS_KERNEL_SPMV) {
        // compulsory bytes: every stored nonzero once, one column index per
        // row (CSR) or per nonzero 3x3 node block (sliced layout), x and y once
        bytes = c->block() ? hb * nnz + (c->col16 ? 2.0 : 4.0) * c->pat.nblocks + 2.0 * hb * n
                           : nnz * (hb + 4.0) + 4.0 * (n + 1) + 2.0 * hb * n;
        if (mixed) {
            fill<float>(c->pu.as<float>(), n, 1.0f, c->stream);
            PcgView<float> v = pcg_view<float>(c, 0.0);
            v.st = bst;
            v.ctl = bctl;
            run([&] { launch_spmv<float>(c, g, v, 0, g.gu, 0); });
        } else {
            fill<double>(c->pu.as<double>(), n, 1.0, c->stream);
            PcgView<double> v = pcg_view<double>(c, 0.0);
            v.st = bst;
            v.ctl = bctl;
            run([&] { launch_spmv<double>(c, g, v, 0, g.gu, 0); });
        }
    } else if (kernel == SORS_KERNEL_PCG_UPDATE) {
        bytes = hb * (11.0 * n + 9.0 * N);
        if (mixed) {
            PcgView<float> v = pcg_view<float>(c, 0.0);
            v.st = bst;
            v.ctl = bctl;
            run([&] { launch_update<float>(c, g, v, 1, g.gu, 0); });
        } else {
            PcgView<double> v = pcg_view<double>(c, 0.0);
            v.st = bst;
            v.ctl = bctl;
            run([&] { launch_update<double>(c, g, v, 1, g.gu, 0); });
        }
    } else if (kernel == SORS_KERNEL_COPY) {
        const long long n4 = nnz / 4;  // int4 elements of the column array
        ensure(c->copybuf, sizeof(int4) * std::max<long long>(n4, 1));
        bytes = 2.0 * 16.0 * n4;
        const int grid = (int)std::min<long long>((n4 + kBlock - 1) / kBlock, 8192);
        run([&] {
            k_copy<<<std::max(grid, 1), kBlock, 0, c->stream>>>(n4, c->cols.as<int4>(), c->copybuf.as<int4>());
        });
    } else if (kernel == SORS_KERNEL_ASSEMBLE) {
        bytes = 4.0 * c->ed * c->E + 16.0 * c->E + 48.0 * N + hb * nnz + 8.0 * n;  // values written once
        ElemArgs a = elem_args(c, c->q.as<double>());
        a.gint = c->gint.as<double>();
        a.vals = c->vals.p;
        a.hscale = 1.0;
        if (mixed)
            run([&] { launch_elem<float, 0>(c, a, c->opt.project_spd); });
        else
            run([&] { launch_elem<double, 0>(c, a, c->opt.project_spd); });
    } else {
        return fail(c, SORS_ERR_ARG, "unknown kernel id");
    }
    if (algo_bytes) *algo_bytes = bytes;
    return SORS_OK;
    SORS_GUARD_END(c)
}

// ---- contact ---------------------------------------------------------------
static bool normalized3(const double* v, double* out) {
    const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (!(l > 0.0) || !std::isfinite(l)) return false;
    for (int k = 0; k < 3; ++k) out[k] = v[k] / l;
    return true;
}

int sors_set_contact(sors_ctx* c, const sors_contact_desc* d) {
    SORS_GUARD(c)
    if (!d) {
        c->has_contact = false;
        c->sqp_built = false;
        return SORS_OK;
    }
    if (d->n_planes < 0 || d->n_planes > kMaxPrim || d->n_disks < 0 || d->n_disks > kMaxPrim)
        return fail(c, SORS_ERR_ARG, "at most 4 planes and 4 disks");
    if ((d->n_planes && (!d->plane_point || !d->plane_normal)) ||
        (d->n_disks && (!d->disk_center || !d->disk_normal || !d->disk_radius)))
        return fail(c, SORS_ERR_ARG, "missing contact primitive arrays");
    if (!(d->stiffness >= 0.0) || !(d->friction >= 0.0) || (d->friction > 0.0 && !(d->eps_v > 0.0)))
        return fail(c, SORS_ERR_ARG, "contact needs stiffness >= 0, friction >= 0 and eps_v > 0 with friction");
    ContactArgs a{};
    a.nplanes = d->n_planes;
    a.ndisks = d->n_disks;
    a.k = d->stiffness;
    a.mu = d->friction;
    a.eps_v = d->eps_v;
    for (int j = 0; j < a.nplanes; ++j) {
        for (int k = 0; k < 3; ++k) a.pp[j][k] = d->plane_point[3 * j + k];
        if (!normalized3(d->plane_normal + 3 * j, a.pn[j])) return fail(c, SORS_ERR_ARG, "zero plane normal");
    }
    for (int j = 0; j < a.ndisks; ++j) {
        for (int k = 0; k < 3; ++k) a.dc[j][k] = a.dc_last[j][k] = d->disk_center[3 * j + k];
        if (!normalized3(d->disk_normal + 3 * j, a.dn[j])) return fail(c, SORS_ERR_ARG, "zero disk normal");
        if (!(d->disk_radius[j] >= 0.0)) return fail(c, SORS_ERR_ARG, "negative disk radius");
        a.dr[j] = d->disk_radius[j];
    }
    std::vector<int> surf;
    if (d->surface_nodes) {
        if (d->n_surface < 0) return fail(c, SORS_ERR_ARG, "negative surface count");
        surf.assign(d->surface_nodes, d->surface_nodes + d->n_surface);
        std::vector<char> seen((size_t)c->N, 0);
        for (int v : surf) {
            if (v < 0 || v >= c->N) return fail(c, SORS_ERR_ARG, "surface vertex index out of range");
            if (seen[(size_t)v]) return fail(c, SORS_ERR_ARG, "surface vertex listed twice");
            seen[(size_t)v] = 1;
        }
    } else {
        surf.resize((size_t)c->N);
        for (int v = 0; v < c->N; ++v) surf[(size_t)v] = v;
    }
    a.nsurf = (int)surf.size();
    upload_vec(c->csurf, surf, c->stream);
    c->h_surf = surf;
    c->sqp_built = false;  // the K table and the disk footprint start afresh
    ensure(c->lagN, sizeof(double) * std::max(a.nsurf, 1));
    ensure(c->lagn, sizeof(double) * 3 * std::max(a.nsurf, 1));
    ensure(c->lagD, sizeof(double) * 3 * std::max(a.nsurf, 1));
    ensure(c->gcon, sizeof(double) * c->n);
    fill<double>(c->gcon.as<double>(), c->n, 0.0, c->stream);
    fill<double>(c->lagN.as<double>(), std::max(a.nsurf, 1), 0.0, c->stream);
    c->cargs = a;
    c->has_contact = (a.nplanes + a.ndisks) > 0;
    contact_lag(c);
    sync(c);
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_set_disk(sors_ctx* c, int idx, const double* center, const double* normal, double radius) {
    SORS_GUARD(c)
    if (!c->has_contact || idx < 0 || idx >= c->cargs.ndisks) return fail(c, SORS_ERR_ARG, "diskIdx out of range in set_disk()");
    if (!center || !normal || !(radius >= 0.0)) return fail(c, SORS_ERR_ARG, "bad disk");
    double nn[3];
    if (!normalized3(normal, nn)) return fail(c, SORS_ERR_ARG, "zero disk normal");
    for (int k = 0; k < 3; ++k) {
        c->cargs.dc[idx][k] = center[k];
        c->cargs.dn[idx][k] = nn[k];
    }
    c->cargs.dr[idx] = radius;
    return SORS_OK;
    SORS_GUARD_END(c)
}

// ---- actuation and adjoint --------------------------------------------------
int sors_set_active(sors_ctx* c, const sors_active_desc* d) {
    SORS_GUARD(c)
    if (!d) {
        c->has_active = false;
        return SORS_OK;
    }
    if (c->ed != 4 || c->material != SORS_NEOHOOKEAN)
        return fail(c, SORS_ERR_UNSUPPORTED, "active fibre stress is implemented for Neo-Hookean tetrahedra");
    if (d->n_groups < 1 || !d->elem_group || !d->fibre || !(d->sigma_max >= 0.0))
        return fail(c, SORS_ERR_ARG, "active stress needs n_groups >= 1, groups, fibres and sigma_max >= 0");
    const int E = c->E;
    std::vector<double> fib((size_t)3 * E);
    for (int e = 0; e < E; ++e) {
        const int gi = d->elem_group[e];
        if (gi < -1 || gi >= d->n_groups) return fail(c, SORS_ERR_ARG, "element muscle group out of range");
        double f[3] = {d->fibre[3 * (size_t)e], d->fibre[3 * (size_t)e + 1], d->fibre[3 * (size_t)e + 2]};
        const double l = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
        if (gi >= 0 && !(l > 0.0)) return fail(c, SORS_ERR_ARG, "zero fibre direction");
        for (int k = 0; k < 3; ++k) fib[(size_t)k * E + e] = l > 0.0 ? f[k] / l : 0.0;
    }
    upload(c->agroup, d->elem_group, (size_t)E, c->stream);
    upload(c->afib, fib.data(), fib.size(), c->stream);
    ensure(c->act, sizeof(double) * d->n_groups);
    fill<double>(c->act.as<double>(), d->n_groups, 0.0, c->stream);
    c->ngroups = d->n_groups;
    c->sigma = d->sigma_max;
    c->has_active = true;
    sync(c);
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_set_activation(sors_ctx* c, const double* act) {
    SORS_GUARD(c)
    if (!c->has_active) return fail(c, SORS_ERR_ARG, "no active stress set (sors_set_active)");
    if (!act) return fail(c, SORS_ERR_ARG, "null activation");
    upload(c->act, act, (size_t)c->ngroups, c->stream);
    sync(c);
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_set_muscle(sors_ctx* c, const sors_muscle_desc* d) {
    SORS_GUARD(c)
    if (!d) {
        c->has_muscle = false;
        return SORS_OK;
    }
    if (c->material == SORS_TEST_POLY)
        return fail(c, SORS_ERR_UNSUPPORTED, "the muscle energy needs a Tetrahedron or Hexahedron element");
    if (d->n_groups < 1 || !d->elem_group || !d->direction || !d->stiffness)
        return fail(c, SORS_ERR_ARG, "muscle needs n_groups >= 1, groups, directions and stiffnesses");
    const int E = c->E;
    std::vector<double> dir((size_t)3 * E);
    for (int e = 0; e < E; ++e) {
        const int gi = d->elem_group[e];
        if (gi < -1 || gi >= d->n_groups) return fail(c, SORS_ERR_ARG, "element muscle group out of range");
        for (int k = 0; k < 3; ++k) dir[(size_t)k * E + e] = d->direction[3 * (size_t)e + k];
    }
    upload(c->mgroup, d->elem_group, (size_t)E, c->stream);
    upload(c->mdir, dir.data(), dir.size(), c->stream);
    upload(c->mk, d->stiffness, (size_t)E, c->stream);
    ensure(c->mact, sizeof(double) * d->n_groups);
    fill<double>(c->mact.as<double>(), d->n_groups, 0.0, c->stream);
    c->mngroups = d->n_groups;
    c->has_muscle = true;
    sync(c);
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_set_muscle_activation(sors_ctx* c, const double* act) {
    SORS_GUARD(c)
    if (!c->has_muscle) return fail(c, SORS_ERR_ARG, "no muscle energy set");
    if (!act) return fail(c, SORS_ERR_ARG, "null activations");
    upload(c->mact, act, (size_t)c->mngroups, c->stream);
    sync(c);
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_set_pressure(sors_ctx* c, const sors_pressure_desc* d) {
    SORS_GUARD(c)
    if (!d || d->n_faces <= 0) {
        c->has_pressure = false;
        c->pnf = 0;
        return SORS_OK;
    }
    const int nv = c->ed == 8 ? 4 : 3;
    if (c->ed != 4 && c->ed != 8) return fail(c, SORS_ERR_ARG, "pressure needs a tet4 or hex8 mesh");
    if (!d->faces || !d->groups) return fail(c, SORS_ERR_ARG, "null faces or groups");
    if (d->n_groups < 1) return fail(c, SORS_ERR_ARG, "n_groups must be >= 1");
    if (!c->block())
        return fail(c, SORS_ERR_UNSUPPORTED, "pressure needs the node-block layout (node-granular Dirichlet mask)");
    const int nf = d->n_faces;
    std::vector<int> soa((size_t)nv * nf), grp(d->groups, d->groups + nf);
    std::vector<int> cnt((size_t)c->N + 1, 0);
    for (int f = 0; f < nf; ++f) {
        if (grp[f] < 0 || grp[f] >= d->n_groups) return fail(c, SORS_ERR_ARG, "surface group out of range");
        for (int j = 0; j < nv; ++j) {
            const int v = d->faces[(size_t)nv * f + j];
            if (v < 0 || v >= c->N) return fail(c, SORS_ERR_ARG, "surface vertex index out of range");
            for (int i = 0; i < j; ++i)
                if (d->faces[(size_t)nv * f + i] == v) return fail(c, SORS_ERR_ARG, "a face lists a vertex twice");
            soa[(size_t)j * nf + f] = v;
            cnt[(size_t)v + 1]++;
        }
    }
    // node -> faces, ascending face index per node
    for (int v = 0; v < c->N; ++v) cnt[(size_t)v + 1] += cnt[v];
    std::vector<int> inc((size_t)cnt[c->N]), pos(cnt.begin(), cnt.end() - 1);
    for (int f = 0; f < nf; ++f)
        for (int j = 0; j < nv; ++j) inc[(size_t)pos[soa[(size_t)j * nf + f]]++] = f;
    upload_vec(c->pfaces, soa, c->stream);
    upload_vec(c->pgroup, grp, c->stream);
    upload_vec(c->pincptr, cnt, c->stream);
    upload_vec(c->pinc, inc, c->stream);
    std::vector<double> zero((size_t)d->n_groups, 0.0);
    upload_vec(c->pact, zero, c->stream);
    ensure(c->pff, sizeof(double) * 3 * (size_t)nf);
    ensure(c->fp0, sizeof(double) * c->n);
    ensure(c->fp1, sizeof(double) * c->n);
    c->pnf = nf;
    c->pnv = nv;
    c->pngroups = d->n_groups;
    c->has_pressure = true;
    sync(c);
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_set_pressure_activation(sors_ctx* c, const double* p) {
    SORS_GUARD(c)
    if (!c->has_pressure) return fail(c, SORS_ERR_ARG, "no pressure force set");
    if (!p) return fail(c, SORS_ERR_ARG, "null pressures");
    upload(c->pact, p, (size_t)c->pngroups, c->stream);
    sync(c);
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_set_youngs(sors_ctx* c, const double* youngs) {
    SORS_GUARD(c)
    if (!youngs) return fail(c, SORS_ERR_ARG, "null Young's moduli");
    std::vector<double> mu(c->E), lam(c->E);
    for (int e = 0; e < c->E; ++e) material_constants(c->material, youngs[e], c->h_nu[e], &mu[e], &lam[e]);
    upload(c->mu, mu.data(), mu.size(), c->stream);
    upload(c->lam, lam.data(), lam.size(), c->stream);
    sync(c);
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_set_state(sors_ctx* c, const double* q, const double* v) {
    SORS_GUARD(c)
    c->vold_valid = false;
    if (q) {
        upload(c->qprev, q, (size_t)c->n, c->stream);
        upload(c->q, q, (size_t)c->n, c->stream);
    }
    if (v) upload(c->vprev, v, (size_t)c->n, c->stream);
    sync(c);
    return SORS_OK;
    SORS_GUARD_END(c)
}

int sors_adjoint_step(sors_ctx* c, const double* q_prev, const double* v_prev, const double* q_next, double dt,
                      const double* qbar_next, const double* vbar_next, double* qbar_prev, double* vbar_prev,
                      double* youngs_bar, double* act_bar, double rtol, int max_iter, long long* pcg_iters) {
    SORS_GUARD(c)
    if (!q_prev || !v_prev || !q_next || !qbar_next || !vbar_next || !qbar_prev || !vbar_prev)
        return fail(c, SORS_ERR_ARG, "null state or adjoint array");
    if (!(dt > 0.0) || max_iter < 1) return fail(c, SORS_ERR_ARG, "dt must be > 0 and max_iter >= 1");
    if (c->scheme != SORS_BACKWARD_EULER || c->method == SORS_MINIMIZE_SQP || c->ed != 4 ||
        c->material != SORS_NEOHOOKEAN)
        return fail(c, SORS_ERR_UNSUPPORTED,
                    "the adjoint is implemented for backward Euler, Newton methods, Neo-Hookean tetrahedra");
    if (c->has_contact && (c->cargs.mu > 0.0 || c->cargs.ndisks > 0))
        return fail(c, SORS_ERR_UNSUPPORTED, "the adjoint takes frictionless plane contact only");
    if (c->has_muscle) return fail(c, SORS_ERR_UNSUPPORTED, "the adjoint does not take the muscle energy");
    if (c->has_pressure) return fail(c, SORS_ERR_UNSUPPORTED, "the adjoint does not take pressure forces");
    const int n = c->n, E = c->E;
    ensure_pattern(c, dt);
    upload(c->qprev, q_prev, (size_t)n, c->stream);
    upload(c->vprev, v_prev, (size_t)n, c->stream);
    upload(c->q, q_next, (size_t)n, c->stream);
    std::vector<double> b((size_t)n);
    for (int i = 0; i < n; ++i) b[(size_t)i] = qbar_next[i] + vbar_next[i] / dt;
    upload(c->adjb, b.data(), b.size(), c->stream);
    ensure(c->adjq, sizeof(double) * n);
    ensure(c->adjv, sizeof(double) * n);
    ensure(c->adjE, sizeof(double) * std::max(E, 1));
    ensure(c->adjA, sizeof(double) * std::max(c->ngroups, 1));
    clear_flag(c);
    assemble<double>(c, c->q.as<double>(), dt, true, 0);
    build_preconditioner<double>(c);
    if (read_flag(c) & 1) return fail(c, SORS_ERR_INVERTED, "Inverted Elements");
    bool brk = false;
    const long long it = pcg<double>(c, c->adjb.as<double>(), 1.0, rtol, max_iter, &brk);
    if (pcg_iters) *pcg_iters = it;
    if (brk) return fail(c, SORS_ERR_NAN, "adjoint: PCG breakdown (Hessian not positive definite)");
    const double* lam = (const double*)c->px.p;
    upload(c->dxd, vbar_next, (size_t)n, c->stream);
    k_adj_state<<<grid_for(n), kBlock, 0, c->stream>>>(n, dt, c->damping, c->mass.as<double>(),
                                                       c->has_mask ? c->mask.as<int>() : nullptr, lam,
                                                       c->dxd.as<double>(), c->adjq.as<double>(),
                                                       c->adjv.as<double>());
    if (E > 0) {
        if (c->nu.bytes < sizeof(double) * (size_t)E) upload(c->nu, c->h_nu.data(), (size_t)E, c->stream);
        ElemArgs a = elem_args(c, c->q.as<double>());
        double* ae = nullptr;
        if (c->has_active) {
            ensure(c->qtmp, sizeof(double) * std::max(E, c->n));
            ae = c->qtmp.as<double>();
        }
        k_adj_params_tet<<<grid_for(E), kBlock, 0, c->stream>>>(a, c->nu.as<double>(), lam, c->adjE.as<double>(), ae);
        if (c->has_active)
            k_group_sum<<<c->ngroups, kBlock, 0, c->stream>>>(E, c->agroup.as<int>(), ae, c->adjA.as<double>());
        else
            fill<double>(c->adjA.as<double>(), 1, 0.0, c->stream);
    }
    HIPX(hipGetLastError());
    HIPX(hipMemcpyAsync(qbar_prev, c->adjq.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPX(hipMemcpyAsync(vbar_prev, c->adjv.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    std::vector<double> eb(youngs_bar ? (size_t)E : 0), ab(act_bar ? (size_t)c->ngroups : 0);
    if (youngs_bar && E > 0)
        HIPX(hipMemcpyAsync(eb.data(), c->adjE.p, sizeof(double) * E, hipMemcpyDeviceToHost, c->stream));
    if (act_bar && c->ngroups > 0)
        HIPX(hipMemcpyAsync(ab.data(), c->adjA.p, sizeof(double) * c->ngroups, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    for (size_t e = 0; e < eb.size(); ++e) youngs_bar[e] += eb[e];
    for (size_t g = 0; g < ab.size(); ++g) act_bar[g] += ab[g];
    return SORS_OK;
    SORS_GUARD_END(c)
}

// ---- serving loop --------------------------------------------------------
void sors_server_default_options(sors_server_options* o) {
    if (!o) return;
    o->hbm_budget_bytes = 200e9;
    o->max_active = 64;
    o->prep_threads = 4;
    o->pcg_chunk = 16;
    sors_default_options(&o->solver);
}

int sors_server_create(const sors_server_options* o, sors_server** out) {
    g_create_error.clear();
    if (!o || !out) {
        g_create_error = "null argument";
        return SORS_ERR_ARG;
    }
    *out = nullptr;
    if (o->max_active < 1 || o->max_active > 256 || o->prep_threads < 1 || o->prep_threads > 64 ||
        o->pcg_chunk < 2 || !(o->hbm_budget_bytes > 0)) {
        g_create_error = "server options out of range (max_active 1..256, prep_threads 1..64, pcg_chunk >= 2)";
        return SORS_ERR_ARG;
    }
    const sors_options& so = o->solver;
    if ((so.precision != SORS_FP64 && so.precision != SORS_MIXED) || so.max_newton < 1 || so.pcg_max_iter < 1 ||
        so.max_backtrack < 0) {
        g_create_error = "bad solver options";
        return SORS_ERR_ARG;
    }
    std::unique_ptr<sors_server> s(new sors_server());
    s->opt = *o;
    try {
        HIPX(hipGetDevice(&s->device));
        HIPX(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        const int M = o->max_active;
        ensure(s->slots, sizeof(SlotRec) * M);
        HIPX(hipMemset(s->slots.p, 0, sizeof(SlotRec) * M));
        // two staging areas (FP32 and FP64 solves may both run in one tick)
        const size_t stage = 2 * stage_area(M);
        static_assert(sizeof(BMember<double>) >= sizeof(BMember<float>), "stage size");
        ensure(s->bmem, stage);
        HIPX(hipHostMalloc(&s->h_stage, stage, hipHostMallocDefault));
        HIPX(hipHostMalloc((void**)&s->h_slots, sizeof(SlotRec) * M, hipHostMallocDefault));
        std::memset(s->h_stage, 0, stage);
        for (int i = M - 1; i >= 0; --i) s->free_slots.push_back(i);
    } catch (const HipError& e) {
        g_create_error = e.msg;
        if (s->h_stage) (void)hipHostFree(s->h_stage);
        if (s->h_slots) (void)hipHostFree(s->h_slots);
        if (s->stream) (void)hipStreamDestroy(s->stream);
        return SORS_ERR_HIP;
    }
    sors_server* sp = s.release();
    for (int i = 0; i < o->prep_threads; ++i) sp->prep.emplace_back(srv_prep_loop, sp);
    sp->worker = std::thread(srv_worker, sp);
    *out = sp;
    return SORS_OK;
}

int sors_server_submit(sors_server* s, const sors_system_desc* d, const double* v0, double dt, int steps,
                       long long* id) {
    if (!s) return SORS_ERR_ARG;
    auto bad = [&](const char* m) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = m;
        return SORS_ERR_ARG;
    };
    if (!d || !id) return bad("null argument");
    if (d->nodes_per_elem != 4 && d->nodes_per_elem != 8) return bad("nodes_per_elem must be 4 or 8");
    if (d->n_nodes <= 0 || d->n_elems <= 0 || !d->rest || !d->conn) return bad("empty or missing mesh arrays");
    if (!(dt > 0.0) || steps < 1) return bad("dt must be > 0 and steps >= 1");
    if (d->method == SORS_MINIMIZE_SQP) return bad("minimize_SQP (OSQP contact) is not implemented on the device");
    if (d->method == SORS_MINIMIZE_FROZENHESSIAN) return bad("the server batches Newton solves: minimize_frozenhessian is not served");
    std::unique_ptr<Request> r(new Request());
    r->N = d->n_nodes;
    r->E = d->n_elems;
    r->ed = d->nodes_per_elem;
    r->material = d->material;
    r->scheme = d->scheme;
    r->method = d->method;
    for (int k = 0; k < 3; ++k) r->grav[k] = d->grav[k];
    r->damping = d->damping_alpha;
    const size_t n = 3 * (size_t)r->N, E = (size_t)r->E;
    r->rest.assign(d->rest, d->rest + n);
    r->conn.assign(d->conn, d->conn + E * r->ed);
    r->youngs.assign(E, 0.0);
    r->poisson.assign(E, 0.0);
    r->density.assign(E, 0.0);
    if (d->youngs) r->youngs.assign(d->youngs, d->youngs + E);
    if (d->poisson) r->poisson.assign(d->poisson, d->poisson + E);
    if (d->density) r->density.assign(d->density, d->density + E);
    if (d->bc_mask) r->bc_mask.assign(d->bc_mask, d->bc_mask + n);
    if (d->bc_value) r->bc_value.assign(d->bc_value, d->bc_value + n);
    if (v0) r->v0.assign(v0, v0 + n);
    r->dt = dt;
    r->steps = steps;
    r->est_bytes = estimate_bytes(r->N, r->E, r->ed, s->opt.solver.coarse_nodes);
    if (r->est_bytes > s->opt.hbm_budget_bytes) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = "request needs more device memory than the server's budget";
        return SORS_ERR_BUDGET;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) {
        s->err = "server is shutting down";
        return SORS_ERR_ARG;
    }
    r->id = s->next_id++;
    r->t_submit = Request::clock::now();
    *id = r->id;
    Request* rp = r.get();
    s->reqs[rp->id] = std::move(r);
    s->pending.push_back(rp);
    s->stats.submitted++;
    srv_try_admit(s);
    return SORS_OK;
}

int sors_server_wait(sors_server* s, long long id, double* q, sors_request_stats* st, double timeout_s) {
    if (!s) return SORS_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) {
        s->err = "unknown request id";
        return SORS_ERR_ARG;
    }
    Request* r = it->second.get();
    auto pred = [&] { return r->done; };
    if (timeout_s < 0)
        s->cv_done.wait(lk, pred);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred))
        return SORS_ERR_TIMEOUT;
    const int rc = r->status;
    if (q && rc == SORS_OK) std::copy(r->q.begin(), r->q.end(), q);
    if (st) *st = r->stats;
    if (rc != SORS_OK) s->err = r->err;
    s->reqs.erase(it);
    return rc;
}

int sors_server_get_stats(sors_server* s, sors_server_stats* st) {
    if (!s || !st) return SORS_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    *st = s->stats;
    return SORS_OK;
}

const char* sors_server_last_error(const sors_server* s) { return s ? s->err.c_str() : g_create_error.c_str(); }

void sors_server_destroy(sors_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (Request* r : s->pending) {
            r->status = SORS_ERR_ARG;
            r->err = "server shut down before the request was admitted";
            r->done = true;
            s->stats.failed++;
        }
        s->pending.clear();
        s->cv_work.notify_all();
        s->cv_prep.notify_all();
        s->cv_done.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->cv_prep.notify_all();
    }
    for (std::thread& t : s->prep)
        if (t.joinable()) t.join();
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->h_stage) (void)hipHostFree(s->h_stage);
    if (s->h_slots) (void)hipHostFree(s->h_slots);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

}  // extern "C"
