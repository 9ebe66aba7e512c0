// coarse.cpp -- coarse space of the two-level preconditioner (host).
#include "coarse.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace sors {

bool build_coarse(int N, const double* rest, const int* mask, const HostPattern& P, int target, CoarseSpace& C,
                  std::string* msg) {
    C = CoarseSpace();
    if (!P.node_granular) {
        if (msg) *msg = "the coarse space needs the node-block layout (node-granular Dirichlet mask)";
        return false;
    }
    auto freenode = [&](int v) { return !mask || mask[3 * v] == 0; };
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    int nfree = 0;
    for (int v = 0; v < N; ++v) {
        if (!freenode(v)) continue;
        ++nfree;
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], rest[3 * v + k]);
            hi[k] = std::max(hi[k], rest[3 * v + k]);
        }
    }
    C.agg.assign(N, -1);
    if (nfree == 0) {
        C.pptr.assign(1, 0);
        C.pchunk.assign(1, 0);
        C.gptr.assign(1, 0);
        C.aptr.assign(1, 0);
        C.dn.assign(4 * (size_t)N, 0.0f);
        const int none = -1;
        for (int v = 0; v < N; ++v) std::memcpy(&C.dn[4 * (size_t)v + 3], &none, sizeof(float));
        return true;
    }
    // bins of equal edge h with about `target` free nodes each
    double ext[3], vol = 1.0;
    int flat = 0;
    const double diag = std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                                  (hi[2] - lo[2]) * (hi[2] - lo[2]));
    for (int k = 0; k < 3; ++k) {
        ext[k] = hi[k] - lo[k];
        if (ext[k] > 1e-12 * std::max(diag, 1e-300))
            vol *= ext[k];
        else
            ++flat;
    }
    const double nb = std::max(1.0, (double)nfree / std::max(target, 1));
    const double h = flat < 3 ? std::pow(vol / nb, 1.0 / (3 - flat)) : 1.0;
    int nbin[3];
    for (int k = 0; k < 3; ++k)
        nbin[k] = (ext[k] > 1e-12 * std::max(diag, 1e-300)) ? std::max(1, (int)std::ceil(ext[k] / h - 1e-9)) : 1;
    std::vector<int> bin_of((size_t)N, -1);
    std::vector<int> count((size_t)nbin[0] * nbin[1] * nbin[2], 0);
    for (int v = 0; v < N; ++v) {
        if (!freenode(v)) continue;
        int b[3];
        for (int k = 0; k < 3; ++k) {
            int t = nbin[k] > 1 ? (int)((rest[3 * v + k] - lo[k]) / ext[k] * nbin[k]) : 0;
            b[k] = std::min(std::max(t, 0), nbin[k] - 1);
        }
        const int id = (b[2] * nbin[1] + b[1]) * nbin[0] + b[0];
        bin_of[v] = id;
        count[id]++;
    }
    std::vector<int> agg_of_bin(count.size(), -1);
    for (size_t b = 0; b < count.size(); ++b)
        if (count[b] > 0) agg_of_bin[b] = C.nagg++;
    for (int v = 0; v < N; ++v)
        if (bin_of[v] >= 0) C.agg[v] = agg_of_bin[bin_of[v]];
    // node lists, centroids, scales
    C.aptr.assign(C.nagg + 1, 0);
    for (int v = 0; v < N; ++v)
        if (C.agg[v] >= 0) C.aptr[C.agg[v] + 1]++;
    for (int a = 0; a < C.nagg; ++a) C.aptr[a + 1] += C.aptr[a];
    C.anodes.assign(C.aptr[C.nagg], 0);
    {
        std::vector<int> fill(C.aptr.begin(), C.aptr.end() - 1);
        for (int v = 0; v < N; ++v)
            if (C.agg[v] >= 0) C.anodes[fill[C.agg[v]]++] = v;
    }
    C.cent.assign(3 * (size_t)C.nagg, 0.0);
    C.scale.assign(C.nagg, 1.0);
    C.rot.assign(C.nagg, 1);
    for (int a = 0; a < C.nagg; ++a) {
        const int m = C.aptr[a + 1] - C.aptr[a];
        double c[3] = {0, 0, 0};
        for (int t = C.aptr[a]; t < C.aptr[a + 1]; ++t)
            for (int k = 0; k < 3; ++k) c[k] += rest[3 * C.anodes[t] + k];
        for (int k = 0; k < 3; ++k) c[k] /= m;
        // second moment of the nodes about the centroid: rotations need it
        // to have rank >= 2 (three non-collinear nodes)
        double S[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, r2 = 0;
        for (int t = C.aptr[a]; t < C.aptr[a + 1]; ++t) {
            double d[3];
            for (int k = 0; k < 3; ++k) d[k] = rest[3 * C.anodes[t] + k] - c[k];
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) S[k][l] += d[k] * d[l];
            r2 += d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        }
        const double s = r2 > 0 ? std::sqrt(r2 / m) : 1.0;
        // rank(S) >= 2 (nodes not all on one line) iff the sum of its principal
        // 2x2 minors is nonzero; tested relative to trace^2
        const double tr = S[0][0] + S[1][1] + S[2][2];
        const double m2 = S[0][0] * S[1][1] - S[0][1] * S[1][0] + S[0][0] * S[2][2] - S[0][2] * S[2][0] +
                          S[1][1] * S[2][2] - S[1][2] * S[2][1];  // sum of principal 2x2 minors
        for (int k = 0; k < 3; ++k) C.cent[3 * (size_t)a + k] = c[k];
        C.scale[a] = s;
        C.rot[a] = (m >= 3 && tr > 0 && m2 > 1e-6 * tr * tr) ? 1 : 0;
    }
    // Galerkin contributions grouped by aggregate pair: row nodes of I, their
    // kept neighbour blocks j with agg(j) = J
    struct Item {
        int J, i, j, base;
    };
    std::vector<std::vector<Item>> rows(C.nagg);
    for (int a = 0; a < C.nagg; ++a) {
        for (int t = C.aptr[a]; t < C.aptr[a + 1]; ++t) {
            const int i = C.anodes[t];
            const int sl = i >> 6, lane = i & 63;
            const int width = (P.sbase[sl + 1] - P.sbase[sl]) / kSlice;
            for (int k = 0; k < width; ++k) {
                const int j = P.bcol[P.sbase[sl] + kSlice * k + lane];
                if (j < 0 || j >= N || C.agg[j] < 0) continue;
                if (j == i && k != P.self_slot[i]) continue;  // padding: the node itself, zero values
                rows[a].push_back({C.agg[j], i, j, (int)block_index(P.sbase.data(), i, k, 0, 0)});
            }
        }
        std::stable_sort(rows[a].begin(), rows[a].end(), [](const Item& x, const Item& y) { return x.J < y.J; });
    }
    C.pptr.assign(1, 0);
    C.pchunk.assign(1, 0);
    C.gptr.assign(1, 0);
    for (int a = 0; a < C.nagg; ++a) {
        const auto& R = rows[a];
        size_t t = 0;
        while (t < R.size()) {
            const int J = R[t].J;
            C.pairI.push_back(a);
            C.pairJ.push_back(J);
            for (; t < R.size() && R[t].J == J; ++t) {
                C.ci.push_back(R[t].i);
                C.cj.push_back(R[t].j);
                C.cbase.push_back(R[t].base);
            }
            const int lo = C.pptr.back(), hi = (int)C.ci.size();
            for (int c = lo; c < hi; c += kGalerkinChunk) C.gptr.push_back(std::min(hi, c + kGalerkinChunk));
            C.pptr.push_back(hi);
            C.pchunk.push_back((int)C.gptr.size() - 1);
        }
    }
    // per-node rigid-body data
    C.dn.assign(4 * (size_t)N, 0.0f);
    for (int v = 0; v < N; ++v) {
        const int a = C.agg[v];
        std::memcpy(&C.dn[4 * (size_t)v + 3], &a, sizeof(float));
        if (a < 0 || !C.rot[a]) continue;
        for (int k = 0; k < 3; ++k)
            C.dn[4 * (size_t)v + k] = (float)((rest[3 * v + k] - C.cent[3 * (size_t)a + k]) / C.scale[a]);
    }
    return true;
}

}  // namespace sors
