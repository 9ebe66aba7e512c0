// host_layout_dump.cpp -- test driver for the host-side layout builders.
//
// Reads a mesh (binary, written by tests/test_host_layout.py), runs
// build_pattern (sliced node-block layout, rank map, slot -> element-block
// lists) and build_coarse (aggregates, rigid-body data, Galerkin pair lists),
// and writes every array to one binary file: a sequence of records
//   name (32 bytes, NUL padded) | element size (int32) | count (int64) | data
// The Python side checks them against brute-force definitions.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "coarse.h"
#include "pattern.h"

namespace {

template <typename T>
void put(FILE* f, const char* name, const std::vector<T>& v) {
    char nm[32] = {0};
    std::strncpy(nm, name, 31);
    const int32_t es = (int32_t)sizeof(T);
    const int64_t n = (int64_t)v.size();
    std::fwrite(nm, 1, 32, f);
    std::fwrite(&es, 4, 1, f);
    std::fwrite(&n, 8, 1, f);
    if (n) std::fwrite(v.data(), sizeof(T), v.size(), f);
}

template <typename T>
bool get(FILE* f, std::vector<T>& v, size_t n) {
    v.resize(n);
    return n == 0 || std::fread(v.data(), sizeof(T), n, f) == n;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s mesh.bin out.bin\n", argv[0]);
        return 2;
    }
    FILE* in = std::fopen(argv[1], "rb");
    if (!in) return 2;
    int32_t hdr[4];  // N, E, ed, coarse target
    if (std::fread(hdr, 4, 4, in) != 4) return 2;
    const int N = hdr[0], E = hdr[1], ed = hdr[2], target = hdr[3];
    std::vector<int> conn, mask;
    std::vector<double> rest;
    if (!get(in, conn, (size_t)E * ed) || !get(in, rest, (size_t)3 * N) || !get(in, mask, (size_t)3 * N)) return 2;
    std::fclose(in);

    sors::HostPattern P;
    std::string msg;
    if (!sors::build_pattern(N, E, ed, conn.data(), mask.data(), true, false, P, &msg)) {
        std::fprintf(stderr, "build_pattern: %s\n", msg.c_str());
        return 1;
    }
    sors::CoarseSpace C;
    if (!sors::build_coarse(N, rest.data(), mask.data(), P, target, C, &msg)) {
        std::fprintf(stderr, "build_coarse: %s\n", msg.c_str());
        return 1;
    }
    FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 2;
    put(out, "sbase", P.sbase);
    put(out, "bcol", P.bcol);
    put(out, "self_slot", P.self_slot);
    put(out, "rank", P.rank);
    put(out, "incptr", P.incptr);
    put(out, "inc", P.inc);
    put(out, "sptr", P.sptr);
    put(out, "sidx", P.sidx);
    put(out, "qslice", P.qslice);
    put(out, "agg", C.agg);
    put(out, "aptr", C.aptr);
    put(out, "anodes", C.anodes);
    put(out, "cent", C.cent);
    put(out, "scale", C.scale);
    put(out, "rot", C.rot);
    put(out, "dn", C.dn);
    put(out, "pairI", C.pairI);
    put(out, "pairJ", C.pairJ);
    put(out, "pptr", C.pptr);
    put(out, "pchunk", C.pchunk);
    put(out, "gptr", C.gptr);
    put(out, "ci", C.ci);
    put(out, "cj", C.cj);
    put(out, "cbase", C.cbase);
    std::fclose(out);
    return 0;
}
