// kernels.h -- gfx950 kernels of the implicit Newton step.
//
// Data layout in HBM (one context = one rollout):
//   conn   int32 [ed][E]        element-major SoA: lane e reads conn[k*E + e]
//   rest   fp64  [3N]           node-major rest positions
//   q, qprev, vprev, g, gint    fp64 [3N] node-major state / residual
//   mu, lam fp64 [E]            Lame constants
//   rowptr int32 [n+1], cols int32 [nnz]     fixed CSR pattern
//   vals   HT [nnz]             HT = float (mixed) or double
//   rank   uint8 [ed*ed][E]     compact scatter map (pattern.h)
//   Minv   HT [N][9]            block-Jacobi inverse
//   PCG vectors HT [n]: r, u, p, s, w, x
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "nh_math.h"
#include "pattern.h"

namespace sors {

constexpr int kBlock = 256;
constexpr int kMaxRedBlocks = 2048;   // partial-sum slots per reduction
constexpr int kNumXcd = 8;            // MI355X: 8 XCDs, blocks dispatched round-robin

enum Scalar {
    S_GG = 0,       // ||g||^2
    S_DD,           // ||dx||^2
    S_GD,           // g . dx
    S_EELEM,        // element energy (absolute or difference)
    S_ENODE,        // node energy terms
    S_AUX,
    S_COUNT
};

// PCG scalars carried between iterations (two copies, used alternately)
struct PcgState {
    double alpha, gamma, rr0, tol2;
    double rr;  // ||r||^2 of iteration iter - 1 (iteration 0 at iter = 0): the host sizes graph chunks from it
    int iter, pad;
};

// PCG outcome, written once by the kernel that ends the solve
struct PcgCtl {
    int done, breakdown, iter, pad;
    double rr;
};

struct Reduce {
    double* partials;    // [NV][kMaxRedBlocks] per-block sums
};

struct ElemArgs {
    int E, ed, N;
    const int* conn;
    const double* rest;
    const double* q;
    const double* dq;     // trial direction or null
    double alpha;         // positions = q + alpha * dq
    const double* mu;
    const double* lam;
    const int* mask;      // 3N or null
    int sqp;
    int material;         // SORS_NEOHOOKEAN | SORS_TEST_POLY | SORS_STABLE_NEOHOOKEAN (mu, lam: mu', lam')
    double hscale;        // element tangent multiplier (BE 1, CN 1/2)
    double* gint;         // fp64 atomics target or null
    void* vals;           // HT atomics target or null
    const int* rowptr;
    const uint8_t* rank;
    const int* fullmap;
    const int* sbase;     // sliced node-block layout of the values (pattern.h) or null for CSR
    int* flag;            // bit 0: inverted element
    double* ex_energy;    // export (test mode) or null
    double* ex_grad;
    double* ex_hess;
    // active fibre stress (tet4): Psi_act = 1/2 sigma_max a (|F f|^2 - 1)
    const int* agroup;    // [E] muscle group or -1; null: no active stress
    const double* afib;   // [3][E] unit fibre direction in the rest state (SoA)
    const double* act;    // [n_groups] activation
    double sigma;         // sigma_max (Pa)
    // SoftCon muscle (softconMuscle.cpp:12-134)
    const int* mgroup;    // [E] muscle group or -1; null: no muscle
    const double* mdir;   // [3][E] direction m (SoA, as given)
    const double* mk;     // [E] stiffness k
    const double* mact;   // [n_muscle_groups] activation
};

struct NodeArgs {
    int n;
    double dt;
    int scheme;
    int sqp;
    double damping;
    double grav[3];
    const double* mass;
    const double* q;
    const double* qprev;
    const double* vprev;
    const double* gprev;
    const double* gint;
    const double* fext;
    const int* mask;
    const double* bcval;
    double* g;            // out
    double* gtmp;         // out: internal gradient (for Crank-Nicolson history)
    void* vals;           // HT or null
    const int* diag;
    const double* gcon;   // contact gradient per DOF (masked DOFs 0) or null
    const double* fpres;  // pressure force at q per DOF or null
    const double* fpres1; // pressure force at q + alpha dq (energy differences) or null
};

// PressureForce faces (pressureForce.cpp): nv = 3 (tet4 meshes) or 4 (hex8)
// vertices per face, SoA faces[j * nf + f]; per-node incident faces in
// ascending order (incptr / inc)
struct PressArgs {
    int nf, nv, N;
    const int* faces;
    const int* group;
    const double* pact;   // pressure per group
    const int* incptr;
    const int* inc;
};

// Contact primitives and per-candidate friction lag (penalty + lagged
// Coulomb friction).  Geometry: plane n.(x - p) (planeContact.cpp:4-42),
// disk n.(x - c) and radial |(x - c) - (n.(x - c)) n| (disk3D.h:33-45).
constexpr int kMaxPrim = 4;
constexpr double kDiskRim = 1e-3;   // rim falloff width (m)
struct ContactArgs {
    int nplanes, ndisks, nsurf, pad;
    double k, mu, eps_v;
    double pp[kMaxPrim][3], pn[kMaxPrim][3];
    double dc[kMaxPrim][3], dn[kMaxPrim][3], dr[kMaxPrim], dc_last[kMaxPrim][3];
    const int* surf;      // candidate nodes
    const int* mask;      // per DOF or null
    double* lagN;         // [nsurf]
    double* lagn;         // [nsurf][3]
    double* lagD;         // [nsurf][3]
};

// host launchers (sors_amd.hip)
template <typename HT>
void launch_assemble(const ElemArgs& a, int project, hipStream_t st);
void launch_element_energy(const ElemArgs& a, int diff, double* out, Reduce red, hipStream_t st);
template <typename HT>
void launch_node_terms(const NodeArgs& a, double* gg_out, Reduce red, hipStream_t st);

}  // namespace sors
