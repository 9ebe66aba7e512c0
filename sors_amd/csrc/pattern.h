// pattern.h -- fixed CSR pattern of the assembled Hessian and the scatter
// map from element-local entries to CSR value slots.
//
// The pattern reproduces what the reference's Energy::compute_hessian
// produces after Eigen's setFromTriplets and the time-integrator diagonal
// (cpp/src/energy.cpp:312-390, cpp/src/integrator.cpp:87-124):
//   {(R,C): DOFs R and C belong to a common element and neither is a
//           Dirichlet DOF (every pair under the SQP semantics)}
//   u {(0,0) if any element pair was skipped -- the unassigned default
//      Triplet() slots of energy.cpp:315 sum into entry (0,0)}
//   u {(i,i) for every DOF if dt != 0}
//   u {(i,i) for every Dirichlet DOF under the Newton methods (:381-388)}.
// Rows are sorted by column; duplicates summed, explicit zeros kept.
//
// Scatter map: entry (e, r, c) of element e's (3*ed)^2 matrix lands at CSR
// value index idx, or nowhere (-1) when the reference skips the triplet
// (energy.cpp:337).  When every node's three DOFs share one Dirichlet flag
// ("node-granular" masks, the case of every reference demo) the map is
// stored compactly as one byte per node pair: the rank of node b among the
// kept neighbours of node a; see rank_index().
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#ifndef SORS_HD
#if defined(__HIPCC__)
#define SORS_HD __host__ __device__ __forceinline__
#else
#define SORS_HD inline
#endif
#endif

namespace sors {

constexpr uint8_t kRankSkip = 255;

// CSR value index of entry (3*na + c, 3*nb + d) for a kept node pair whose
// rank is `rank`: all three rows of node na have the same columns, and the
// three columns of node nb are contiguous.
SORS_HD long long rank_index(const int* rowptr, int na, int c, int rank, int d) {
    long long base = rowptr[3 * na];
    long long len = rowptr[3 * na + 1] - base;
    return base + c * len + 3 * rank + d;
}

// Sliced node-block layout of the same matrix (used for the device values
// whenever the map is node-granular).  Nodes are grouped in slices of 64
// (one wavefront, one lane per node = three matrix rows).  Node v lists its
// kept neighbour nodes in ascending order -- slot k of v is the node of rank
// k, so the rank map addresses slots directly -- and a Dirichlet node lists
// only itself (its rows are the identity).  Slice s holds width_s slots per
// lane, padded with the lane's own node and zero values:
//   column of (v, slot k):      bcol[sbase[s] + 64 k + lane]
//   value (v, slot k, c, d):    vals[9 sbase[s] + 64 (9 k + 3 c + d) + lane]
// with s = v / 64, lane = v % 64.
constexpr int kSlice = 64;

SORS_HD long long block_index(const int* sbase, int node, int slot, int c, int d) {
    return 9LL * sbase[node >> 6] + (long long)kSlice * (9 * slot + 3 * c + d) + (node & (kSlice - 1));
}

struct HostPattern {
    int n = 0;                      // DOFs
    long long nnz = 0;
    bool dt_nonzero = true;
    bool node_granular = true;      // compact rank map in use
    std::vector<int> rowptr;        // n + 1
    std::vector<int> cols;          // nnz
    std::vector<int> diag;          // n: index of (i,i) in the CSR values
    std::vector<uint8_t> rank;      // [ed*ed][E] when node_granular
    std::vector<int> fullmap;       // [E][(3ed)^2] otherwise
    // sliced node-block layout (node_granular only)
    int nslices = 0;
    std::vector<int> sbase;         // nslices + 1, prefix sums of 64 * width
    std::vector<int> bcol;          // sbase[nslices]
    std::vector<int> self_slot;     // N
    std::vector<int> bdiag;         // n: block-layout index of (i,i)
    long long nblocks = 0;          // 3x3 node blocks holding at least one CSR entry
    std::vector<int> incptr;        // N + 1: node -> incident (element, local node) list
    std::vector<int> inc;           // e * ed + local index, element order per node
    // tet4 / hex8: element blocks summed into each slot of the sliced layout
    // (empty when the largest entry exceeds int32).  Slot position p =
    // sbase[s] + 64 k + lane (the index of bcol) receives blocks
    // sidx[sptr[p] .. sptr[p+1]), in ascending element order, each naming
    // element e's block of local nodes (la, lb):
    //   hex8: elem_row(e, ed la + lb, ed^2);
    //   tet4: 2 elem_row(e, t, 10) + f, where t = 4 a - a (a - 1) / 2 + b - a
    //   indexes (a, b) = (min, max)(la, lb) in the row-major upper triangle
    //   (the ten stored blocks of a symmetric element Hessian) and f = 1 when
    //   la > lb (the stored block is read transposed).
    std::vector<int> sptr;          // sbase[nslices] + 1
    std::vector<int> sidx;
    std::vector<int> qslice;        // sbase[nslices] / 64: slice of each group of 64 slot positions
};

// Element rows between the two assembly phases.  tet4 (R = 4 gradient rows,
// R = 10 Hessian blocks): element e's rows are contiguous, e R + j, so a slot
// lane that gathers one of its blocks pulls in cache lines its neighbours'
// lanes gather from the same element (assembly 1.25 -> 1.14 ms on the
// 1.92 M-tet beam).  hex8 (R = 8, 64): rows are tiled by 64 elements (one
// wavefront), ((e / 64) R + j) 64 + e % 64, so the eight lanes of an element
// store coalesced (element-contiguous hex rows measured 1.20 -> 1.30 ms).
// (constexpr: also callable from device code.)
constexpr long long elem_row(long long e, int j, int R) {
    return (R == 4 || R == 10) ? e * R + j : ((e >> 6) * R + j) * 64 + (e & 63);
}
constexpr long long elem_rows_padded(long long E) { return (E + 63) / 64 * 64; }

// mask: 3N flags (may be null).  sqp: SQP assembly semantics (no masking).
// Returns false (with msg) on invalid connectivity.
bool build_pattern(int n_nodes, int n_elems, int ed, const int* conn, const int* mask, bool dt_nonzero, bool sqp,
                   HostPattern& out, std::string* msg);

// full (e, r, c) -> index map from a built pattern (for export / checks)
void scatter_map_full(const HostPattern& p, int n_elems, int ed, const int* conn, const int* mask, bool sqp,
                      int* map);

}  // namespace sors
