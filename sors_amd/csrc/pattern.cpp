// pattern.cpp -- host construction of the fixed CSR pattern and scatter map
// (rule and layout documented in pattern.h).
#include <string>

#include "pattern.h"

#include <algorithm>
#include <cstring>
#include <thread>

namespace sors {

namespace {

// run f(lo, hi) over [0, n) on up to `threads` host threads
template <typename F>
void parallel_for(long long n, F f) {
    unsigned hw = std::thread::hardware_concurrency();
    int threads = (int)std::min<unsigned>(hw ? hw : 1, 16);
    if (n < 20000 || threads <= 1) {
        f(0LL, n);
        return;
    }
    std::vector<std::thread> ts;
    long long chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        long long lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo >= hi) break;
        ts.emplace_back([=] { f(lo, hi); });
    }
    for (auto& t : ts) t.join();
}

}  // namespace

bool build_pattern(int N, int E, int ed, const int* conn, const int* mask, bool dtnz, bool sqp, HostPattern& P,
                   std::string* msg) {
    const int n = 3 * N;
    P = HostPattern();
    P.n = n;
    P.dt_nonzero = dtnz;
    auto masked = [&](int dof) { return !sqp && mask && mask[dof] != 0; };
    if ((long long)E * ed > 0x7fffffffLL) {
        if (msg) *msg = "mesh exceeds 2^31 element-node incidences";
        return false;
    }
    for (long long i = 0; i < (long long)E * ed; ++i)
        if (conn[i] < 0 || conn[i] >= N) {
            if (msg) *msg = "element index out of range";
            return false;
        }
    // node granularity of the Dirichlet mask
    bool gran = true;
    if (mask && !sqp)
        for (int v = 0; v < N && gran; ++v)
            if ((mask[3 * v] != 0) != (mask[3 * v + 1] != 0) || (mask[3 * v] != 0) != (mask[3 * v + 2] != 0))
                gran = false;
    // any skipped element pair -> default triplet at (0,0)
    bool anyskip = false;
    if (mask && !sqp)
        for (long long i = 0; i < (long long)E * ed && !anyskip; ++i)
            for (int k = 0; k < 3; ++k)
                if (mask[3 * conn[i] + k]) anyskip = true;

    // node -> incident elements
    std::vector<int> iptr(N + 1, 0), inc((size_t)E * ed);
    for (long long i = 0; i < (long long)E * ed; ++i) iptr[conn[i] + 1]++;
    for (int v = 0; v < N; ++v) iptr[v + 1] += iptr[v];
    {
        std::vector<int> fill(iptr.begin(), iptr.end() - 1);
        for (int e = 0; e < E; ++e)
            for (int a = 0; a < ed; ++a) inc[fill[conn[(size_t)e * ed + a]]++] = e * ed + a;
    }
    // node -> sorted unique neighbour nodes (self included)
    std::vector<int> nptr(N + 1, 0);
    std::vector<std::vector<int>> nbr(N);
    parallel_for(N, [&](long long lo, long long hi) {
        std::vector<int> tmp;
        for (long long v = lo; v < hi; ++v) {
            tmp.clear();
            for (int t = iptr[v]; t < iptr[v + 1]; ++t)
                for (int a = 0; a < ed; ++a) tmp.push_back(conn[(size_t)(inc[t] / ed) * ed + a]);
            std::sort(tmp.begin(), tmp.end());
            tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
            nbr[v] = tmp;
        }
    });
    // row lengths
    P.rowptr.assign(n + 1, 0);
    auto forced = [&](int R) { return dtnz || masked(R) || (R == 0 && anyskip); };
    parallel_for(n, [&](long long lo, long long hi) {
        for (long long R = lo; R < hi; ++R) {
            int v = (int)(R / 3), m = 0;
            bool hasdiag = false;
            if (!masked((int)R))
                for (int u : nbr[v])
                    for (int l = 0; l < 3; ++l)
                        if (!masked(3 * u + l)) {
                            ++m;
                            if (3 * u + l == R) hasdiag = true;
                        }
            if (!hasdiag && forced((int)R)) ++m;
            P.rowptr[R + 1] = m;
        }
    });
    for (int R = 0; R < n; ++R) P.rowptr[R + 1] += P.rowptr[R];
    P.nnz = P.rowptr[n];
    if (P.nnz > 0x7fffffffLL) {
        if (msg) *msg = "pattern exceeds 2^31 entries";
        return false;
    }
    P.cols.resize((size_t)P.nnz);
    P.diag.assign(n, -1);
    parallel_for(n, [&](long long lo, long long hi) {
        for (long long R = lo; R < hi; ++R) {
            int v = (int)(R / 3);
            int* out = P.cols.data() + P.rowptr[R];
            int m = 0;
            bool need = forced((int)R);
            bool placed = false;
            if (!masked((int)R))
                for (int u : nbr[v])
                    for (int l = 0; l < 3; ++l) {
                        int C = 3 * u + l;
                        if (masked(C)) continue;
                        if (need && !placed && C > R) {
                            out[m++] = (int)R;
                            placed = true;
                        }
                        if (C == R) placed = true;
                        out[m++] = C;
                    }
            if (need && !placed) out[m++] = (int)R;
            for (int k = 0; k < m; ++k)
                if (out[k] == R) P.diag[R] = P.rowptr[R] + k;
        }
    });
    // scatter map
    P.node_granular = gran;
    if (gran) {
        bool fits = true;
        for (int v = 0; v < N && fits; ++v)
            if (nbr[v].size() >= kRankSkip) fits = false;
        P.node_granular = fits;
    }
    if (P.node_granular) {
        P.rank.assign((size_t)ed * ed * E, kRankSkip);
        parallel_for(E, [&](long long lo, long long hi) {
            std::vector<int> kept;
            for (long long e = lo; e < hi; ++e)
                for (int a = 0; a < ed; ++a) {
                    int na = conn[e * ed + a];
                    if (masked(3 * na)) continue;
                    for (int b = 0; b < ed; ++b) {
                        int nb = conn[e * ed + b];
                        if (masked(3 * nb)) continue;
                        // rank of nb among the kept neighbours of na
                        int r = 0;
                        for (int u : nbr[na]) {
                            if (u == nb) break;
                            if (!masked(3 * u)) ++r;
                        }
                        P.rank[(size_t)(a * ed + b) * E + e] = (uint8_t)r;
                    }
                }
        });
        // sliced node-block layout
        const int ns = (N + kSlice - 1) / kSlice;
        P.nslices = ns;
        P.sbase.assign(ns + 1, 0);
        P.self_slot.assign(N, 0);
        std::vector<int> deg(N, 1);
        long long nblocks = 0;
        for (int v = 0; v < N; ++v) {
            if (masked(3 * v)) {
                deg[v] = 1;
                ++nblocks;
                continue;
            }
            int k = 0, self = -1;
            for (int u : nbr[v]) {
                if (masked(3 * u)) continue;
                if (u == v) self = k;
                ++k;
            }
            if (k == 0) {
                // isolated node: only its (forced) diagonal can exist
                deg[v] = 1;
                self = 0;
                nblocks += forced(3 * v) ? 1 : 0;
            } else {
                deg[v] = k;
                nblocks += k;
            }
            P.self_slot[v] = self < 0 ? 0 : self;
        }
        P.nblocks = nblocks;
        for (int sl = 0; sl < ns; ++sl) {
            int w = 1;
            for (int v = sl * kSlice; v < std::min(N, (sl + 1) * kSlice); ++v) w = std::max(w, deg[v]);
            P.sbase[sl + 1] = P.sbase[sl] + w * kSlice;
        }
        if (9LL * P.sbase[ns] > 0x7fffffffLL) {
            if (msg) *msg = "block layout exceeds 2^31 values";
            return false;
        }
        P.bcol.assign(P.sbase[ns], 0);
        parallel_for(ns, [&](long long lo, long long hi) {
            for (long long sl = lo; sl < hi; ++sl) {
                const int w = (P.sbase[sl + 1] - P.sbase[sl]) / kSlice;
                for (int lane = 0; lane < kSlice; ++lane) {
                    const int v = (int)sl * kSlice + lane;
                    int* col = P.bcol.data() + P.sbase[sl] + lane;
                    int k = 0;
                    if (v < N && !masked(3 * v))
                        for (int u : nbr[v])
                            if (!masked(3 * u)) col[kSlice * k++] = u;
                    for (; k < w; ++k) col[kSlice * k] = (v < N) ? v : 0;  // padding: own node, zero values
                }
            }
        });
        P.bdiag.assign(n, -1);
        for (int v = 0; v < N; ++v)
            for (int c = 0; c < 3; ++c) P.bdiag[3 * v + c] = (int)block_index(P.sbase.data(), v, P.self_slot[v], c, c);
        // node -> (element, local index) incidence for the per-node gradient sums
        P.incptr = std::move(iptr);
        P.inc = std::move(inc);
        // largest list entry: 2 (10 E') + 1 for tet4, 64 E' for hex8 (E' = E padded to 64)
        const long long top = (ed == 4 ? 20LL : 64LL) * elem_rows_padded(E);
        if ((ed == 4 || ed == 8) && top <= 0x7fffffffLL) {
            // slot -> element blocks (counting sort by slot position; elements ascending)
            const long long S = P.sbase[ns];
            P.qslice.assign((size_t)(S / kSlice), 0);
            for (int sl = 0; sl < ns; ++sl)
                for (int q = P.sbase[sl] / kSlice; q < P.sbase[sl + 1] / kSlice; ++q) P.qslice[q] = sl;
            P.sptr.assign((size_t)S + 1, 0);
            auto slot_of = [&](int e, int la, int lb) -> long long {
                const int rk = P.rank[(size_t)(la * ed + lb) * E + e];
                if (rk == kRankSkip) return -1;
                const int v = conn[(size_t)e * ed + la];
                return (long long)P.sbase[v >> 6] + (long long)kSlice * rk + (v & (kSlice - 1));
            };
            for (int e = 0; e < E; ++e)
                for (int la = 0; la < ed; ++la)
                    for (int lb = 0; lb < ed; ++lb) {
                        const long long p = slot_of(e, la, lb);
                        if (p >= 0) P.sptr[p + 1]++;
                    }
            for (long long p = 0; p < S; ++p) P.sptr[p + 1] += P.sptr[p];
            P.sidx.assign((size_t)P.sptr[S], 0);
            std::vector<int> fill(P.sptr.begin(), P.sptr.end() - 1);
            for (int e = 0; e < E; ++e)
                for (int la = 0; la < ed; ++la)
                    for (int lb = 0; lb < ed; ++lb) {
                        const long long p = slot_of(e, la, lb);
                        if (p < 0) continue;
                        if (ed == 4) {
                            // upper-triangle block t(a, b) of element e, transposed when la > lb
                            const int a = std::min(la, lb), b = std::max(la, lb);
                            const int t = 4 * a - a * (a - 1) / 2 + b - a;
                            P.sidx[fill[p]++] = (int)(2 * elem_row(e, t, 10) + (la > lb ? 1 : 0));
                        } else {
                            P.sidx[fill[p]++] = (int)elem_row(e, ed * la + lb, ed * ed);
                        }
                    }
        }
    } else {
        int nd = 3 * ed;
        P.fullmap.resize((size_t)E * nd * nd);
        scatter_map_full(P, E, ed, conn, mask, sqp, P.fullmap.data());
    }
    return true;
}

void scatter_map_full(const HostPattern& P, int E, int ed, const int* conn, const int* mask, bool sqp, int* map) {
    const int nd = 3 * ed;
    auto masked = [&](int dof) { return !sqp && mask && mask[dof] != 0; };
    parallel_for(E, [&](long long lo, long long hi) {
        for (long long e = lo; e < hi; ++e)
            for (int r = 0; r < nd; ++r) {
                int na = conn[e * ed + r / 3], c = r % 3;
                int R = 3 * na + c;
                for (int cc = 0; cc < nd; ++cc) {
                    int nb = conn[e * ed + cc / 3], d = cc % 3;
                    int C = 3 * nb + d;
                    int* slot = map + ((size_t)e * nd + r) * nd + cc;
                    if (masked(R) || masked(C)) {
                        *slot = -1;
                        continue;
                    }
                    if (P.node_granular) {
                        int rk = P.rank[(size_t)((r / 3) * ed + cc / 3) * E + e];
                        *slot = (int)rank_index(P.rowptr.data(), na, c, rk, d);
                    } else {
                        const int* b = P.cols.data() + P.rowptr[R];
                        const int* en = P.cols.data() + P.rowptr[R + 1];
                        const int* it = std::lower_bound(b, en, C);
                        *slot = (it != en && *it == C) ? (int)(it - P.cols.data()) : -1;
                    }
                }
            }
    });
}

}  // namespace sors
