// coarse.h -- host set-up of the two-level preconditioner's coarse space.
//
// The fine matrix lives in the sliced 3x3 node-block layout (pattern.h).
// Free nodes are grouped into spatial aggregates (bins of the rest-state
// bounding box).  Aggregate I spans the six rigid-body modes of its nodes,
//   Z_i = [ I_3 | -[(x_i - c_I)/s_I]_x ]    (3 x 6 per node i of I),
// translations and rotations about its centroid c_I scaled by its radius
// s_I; an aggregate whose nodes cannot carry rotations (fewer than three
// non-collinear nodes) keeps translations only.  The coarse operator is the
// dense Galerkin product E = Z^T A Z (6 unknowns per aggregate).  For its
// assembly the fine blocks are grouped by aggregate pair (I, J): pair p owns
// contributions [pptr[p], pptr[p+1]) of (i, j, value base) triples, so one
// wavefront per pair sums its 6x6 block in a fixed order.
#pragma once

#include <string>
#include <vector>

#include "pattern.h"

namespace sors {

struct CoarseSpace {
    int nagg = 0;                  // aggregates (6 coarse unknowns each)
    std::vector<int> agg;          // [N] aggregate of a node, -1 for Dirichlet nodes
    std::vector<int> aptr, anodes; // aggregate -> node list (CSR)
    std::vector<double> cent;      // [nagg][3] centroid of the rest positions
    std::vector<double> scale;     // [nagg] radius used to scale the rotations
    std::vector<int> rot;          // [nagg] 1: rotations kept, 0: translations only
    // [N][4] per node: d = (x - c)/s as FP32 (zero where the aggregate keeps
    // no rotations), then the aggregate index bit-cast to float (-1: Dirichlet).
    // The device applies Z from these values only, so restriction, prolongation
    // and the Galerkin product use one and the same Z.
    std::vector<float> dn;
    // Galerkin product, grouped by aggregate pair, each pair's contributions
    // cut into chunks of at most kGalerkinChunk
    std::vector<int> pairI, pairJ; // [npairs]
    std::vector<int> pptr;         // npairs + 1, into the contributions
    std::vector<int> pchunk;       // npairs + 1, into the chunks
    std::vector<int> gptr;         // ngchunks + 1, into the contributions
    std::vector<int> ci, cj;       // [ncontrib] fine row node, fine column node
    std::vector<int> cbase;        // [ncontrib] value index of the block's (0,0) entry
};

constexpr int kGalerkinChunk = 256;   // fine blocks summed by one wavefront

// target: nodes per aggregate (the bins are sized for it).  Needs the sliced
// node-block layout (P.node_granular).
bool build_coarse(int N, const double* rest, const int* mask, const HostPattern& P, int target, CoarseSpace& C,
                  std::string* msg);

}  // namespace sors
