// sors_io.cpp -- host-side IO of the reference (cpp/src/io.cpp) without MshIO
// or vtu11; see sors_io.h.
#include "sors_io.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <map>
#include <queue>
#include <set>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

namespace sors_host {

// ---------------------------------------------------------------- Plane3D
namespace {
constexpr double kPlaneEps = 1e-6;  // EPSILON_DP (common.h:37)
using V3 = std::array<double, 3>;
V3 sub3(const V3& a, const V3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
V3 cross3(const V3& a, const V3& b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
double dot3(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
double norm3(const V3& a) { return std::sqrt(dot3(a, a)); }
}  // namespace

// plane3D.h:18: the default plane (z = 0, normal +z); its square is built as
// for any other plane (the reference leaves it empty, which would give a VTU
// with dangling indices)
Plane3D::Plane3D() { squarePoints_ = create_square_points(1.0); }

// plane3D.h:21-22: the normal is normalised
Plane3D::Plane3D(const V3& point, const V3& normal) : point_(point) {
    const double l = norm3(normal);
    if (!(l > 0.0)) throw std::invalid_argument("plane normal must be nonzero");
    normal_ = {normal[0] / l, normal[1] / l, normal[2] / l};
    squarePoints_ = create_square_points(1.0);
}

bool Plane3D::is_point_on_plane(const V3& x) const { return std::fabs(dot3(sub3(x, point_), normal_)) < kPlaneEps; }

double Plane3D::distance_to_point(const V3& x) const { return dot3(sub3(x, point_), normal_); }

V3 Plane3D::project_point_on_plane(const V3& x) const {
    const double d = distance_to_point(x);
    return {x[0] - d * normal_[0], x[1] - d * normal_[1], x[2] - d * normal_[2]};
}

// plane3D.h:44-63: u = x_hat x n (y_hat x n when n is x_hat), v = n x u, the
// square p +- s/2 u +- s/2 v in the order (+,+) (-,+) (-,-) (+,-).  When n is
// -x_hat, x_hat x n vanishes; y_hat x n is used there too (the reference
// would normalise a zero vector)
std::vector<double> Plane3D::create_square_points(double s) const {
    const V3 x{1.0, 0.0, 0.0}, y{0.0, 1.0, 0.0};
    V3 u = cross3(x, normal_);
    const double dx = norm3(sub3(normal_, x));
    if (dx <= 1e-12 * std::min(norm3(normal_), 1.0) || norm3(u) == 0.0) u = cross3(y, normal_);
    V3 v = cross3(normal_, u);
    const double lu = norm3(u), lv = norm3(v);
    for (int k = 0; k < 3; ++k) {
        u[k] /= lu;
        v[k] /= lv;
    }
    std::vector<double> out;
    out.reserve(12);
    const double su[4] = {0.5, -0.5, -0.5, 0.5}, sv[4] = {0.5, 0.5, -0.5, -0.5};
    for (int c = 0; c < 4; ++c)
        for (int k = 0; k < 3; ++k) out.push_back(point_[k] + su[c] * s * u[k] + sv[c] * s * v[k]);
    return out;
}

void Plane3D::set_point(const V3& p) {
    point_ = p;
    squarePoints_ = create_square_points(1.0);
}

// plane3D.h:73-76: stored as given
void Plane3D::set_normal(const V3& n) {
    normal_ = n;
    squarePoints_ = create_square_points(1.0);
}

// ---------------------------------------------------------------- VTU / PVD
namespace {

struct DataSet {
    std::string name;
    bool point;  // PointData or CellData
    int ncomp;
    std::vector<double> values;
};

void write_array(std::ostream& f, const char* type, const std::string& name, int ncomp,
                 const std::function<void(std::ostream&)>& body) {
    f << "        <DataArray type=\"" << type << "\"";
    if (!name.empty()) f << " Name=\"" << name << "\"";
    if (ncomp > 0) f << " NumberOfComponents=\"" << ncomp << "\"";
    f << " format=\"ascii\">\n";
    body(f);
    f << "\n        </DataArray>\n";
}

// io.cpp:3-175 (tet4, VTK cell 10) and 178-336 (hex8, VTK cell 12): the mesh,
// one quad (VTK cell 9) per plane with its square points, the caller's
// datasets (values at or below a dataset's threshold -- for vectors the norm
// -- written as zero; plane points and cells padded with zeros), then the
// default datasets Displacement (x + y + z per vertex) and
// defaultElementColor (1 per element).
bool save_vtu(const std::string& filename, const Mat& V, const std::vector<int>& elements, int ed, int vtk_type,
              std::vector<VisualOption> visuals, const std::vector<Plane3D>& planes) {
    if (V.cols != 3) throw std::invalid_argument("vertices must be N x 3");
    if (elements.size() % ed != 0) throw std::invalid_argument("elements must be E x " + std::to_string(ed));
    const long nv = V.rows, ne = (long)(elements.size() / ed);
    const long np = (long)planes.size();
    const char* what = ed == 4 ? "Tet" : "Hex";
    VisualOption disp;
    disp.name = "Displacement";
    disp.type = VisualType::ScalarVertex;
    disp.data = Mat(1, nv);
    for (long i = 0; i < nv; ++i) disp.data(0, i) = V(i, 0) + V(i, 1) + V(i, 2);
    visuals.push_back(disp);
    VisualOption color;
    color.name = "defaultElementColor";
    color.type = VisualType::ScalarElement;
    color.data = Mat(1, ne, 1.0);
    visuals.push_back(color);
    std::vector<DataSet> sets;
    for (const auto& vo : visuals) {
        const bool point = vo.type == VisualType::ScalarVertex || vo.type == VisualType::VectorVertex;
        const bool vec = vo.type == VisualType::VectorVertex || vo.type == VisualType::VectorElement;
        const long cnt = point ? nv : ne;
        const int nc = vec ? 3 : 1;
        if (vo.data.rows != nc || vo.data.cols != cnt) {
            static const char* kind[] = {"VectorVertex", "VectorElement", "ScalarVertex", "ScalarElement"};
            throw std::runtime_error(vo.name + " (" + kind[(int)vo.type] + ") must be " + std::to_string(nc) +
                                     (point ? " x numVertices" : " x numElements") + what);
        }
        DataSet ds{vo.name, point, nc, {}};
        ds.values.reserve((size_t)nc * (cnt + (point ? 4 * np : np)));
        for (long i = 0; i < cnt; ++i) {
            if (!vec) {
                const double x = vo.data(0, i);
                ds.values.push_back(vo.threshold < x ? x : 0.0);
            } else {
                const double x = vo.data(0, i), y = vo.data(1, i), z = vo.data(2, i);
                const bool keep = vo.threshold < std::sqrt(x * x + y * y + z * z);
                ds.values.push_back(keep ? x : 0.0);
                ds.values.push_back(keep ? y : 0.0);
                ds.values.push_back(keep ? z : 0.0);
            }
        }
        ds.values.insert(ds.values.end(), (size_t)nc * (point ? 4 * np : np), 0.0);
        sets.push_back(std::move(ds));
    }
    std::ofstream f(filename + ".vtu");
    if (!f) throw std::runtime_error("cannot write " + filename + ".vtu");
    f << std::setprecision(17);
    f << "<?xml version=\"1.0\"?>\n";
    f << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
    f << "  <UnstructuredGrid>\n";
    f << "    <Piece NumberOfPoints=\"" << nv + 4 * np << "\" NumberOfCells=\"" << ne + np << "\">\n";
    for (int pass = 0; pass < 2; ++pass) {
        const bool point = pass == 0;
        f << (point ? "      <PointData>\n" : "      <CellData>\n");
        for (const auto& ds : sets)
            if (ds.point == point)
                write_array(f, "Float64", ds.name, ds.ncomp, [&](std::ostream& o) {
                    for (size_t i = 0; i < ds.values.size(); ++i) o << (i ? " " : "") << ds.values[i];
                });
        f << (point ? "      </PointData>\n" : "      </CellData>\n");
    }
    f << "      <Points>\n";
    write_array(f, "Float64", "", 3, [&](std::ostream& o) {
        for (long i = 0; i < nv; ++i) o << (i ? " " : "") << V(i, 0) << " " << V(i, 1) << " " << V(i, 2);
        for (const auto& pl : planes)
            for (double c : pl.get_square_points()) o << " " << c;
    });
    f << "      </Points>\n      <Cells>\n";
    write_array(f, "Int64", "connectivity", 0, [&](std::ostream& o) {
        for (size_t i = 0; i < elements.size(); ++i) o << (i ? " " : "") << elements[i];
        for (long p = 0; p < np; ++p)
            for (int c = 0; c < 4; ++c) o << " " << nv + 4 * p + c;
    });
    write_array(f, "Int64", "offsets", 0, [&](std::ostream& o) {
        long off = 0;
        for (long e = 0; e < ne; ++e) o << (e ? " " : "") << (off += ed);
        for (long p = 0; p < np; ++p) o << " " << (off += 4);
    });
    write_array(f, "UInt8", "types", 0, [&](std::ostream& o) {
        for (long e = 0; e < ne; ++e) o << (e ? " " : "") << vtk_type;
        for (long p = 0; p < np; ++p) o << " " << 9;
    });
    f << "      </Cells>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n";
    return (bool)f;
}

}  // namespace

bool IO::save_tet_VTU(const std::string& filename, const Mat& vertices, const std::vector<int>& elements,
                      std::vector<VisualOption> visuals, const std::vector<Plane3D>& planes) {
    return save_vtu(filename, vertices, elements, 4, 10, std::move(visuals), planes);
}

bool IO::save_hex_VTU(const std::string& filename, const Mat& vertices, const std::vector<int>& elements,
                      std::vector<VisualOption> visuals, const std::vector<Plane3D>& planes) {
    return save_vtu(filename, vertices, elements, 8, 12, std::move(visuals), planes);
}

// io.cpp:339-356
bool IO::save_pvd(const std::string& filename, const std::string& outputFolder, int timesteps, double dt) {
    std::ofstream f(outputFolder + "/" + filename + ".pvd");
    if (!f) throw std::runtime_error("cannot write " + outputFolder + "/" + filename + ".pvd");
    f << "<?xml version='1.0' encoding='UTF-8'?>\n";
    f << "<VTKFile type='Collection' version='0.1' byte_order='LittleEndian' compressor='vtkZLibDataCompressor'>\n";
    f << "\t<Collection>\n";
    for (int i = 0; i < timesteps; i++)
        f << "\t\t<DataSet timestep='" << i * dt << "' group='' part='0' file='" << filename << "_" << i
          << ".vtu'/>\n";
    f << "\t</Collection>\n";
    f << "</VTKFile>";
    return (bool)f;
}

// ---------------------------------------------------------------- Gmsh .msh
namespace {

// nodes per element of the Gmsh element types 1-31
int gmsh_nodes(int type) {
    static const int n[32] = {0, 2, 3, 4, 4, 8, 6, 5, 3, 6, 9, 10, 27, 18, 14, 1, 8, 20, 15, 13, 9, 10, 12, 15, 15,
                              21, 4, 5, 6, 20, 35, 56};
    if (type < 1 || type > 31) throw std::runtime_error("unsupported Gmsh element type " + std::to_string(type));
    return n[type];
}

struct MshReader {
    std::string buf;
    size_t pos = 0;
    bool binary = false;
    int version = 0;  // 22 or 41
    std::string file;

    [[noreturn]] void fail(const std::string& m) const { throw std::runtime_error(file + ": " + m); }
    bool eof() const { return pos >= buf.size(); }
    std::string line() {
        const size_t e = buf.find('\n', pos);
        std::string s = buf.substr(pos, e == std::string::npos ? std::string::npos : e - pos);
        pos = e == std::string::npos ? buf.size() : e + 1;
        while (!s.empty() && (s.back() == '\r' || s.back() == ' ' || s.back() == '\t')) s.pop_back();
        size_t b = 0;
        while (b < s.size() && (s[b] == ' ' || s[b] == '\t')) ++b;
        return s.substr(b);
    }
    void skip_ws() {
        while (pos < buf.size() && (buf[pos] == ' ' || buf[pos] == '\n' || buf[pos] == '\r' || buf[pos] == '\t'))
            ++pos;
    }
    std::string token() {
        skip_ws();
        const size_t b = pos;
        while (pos < buf.size() && !(buf[pos] == ' ' || buf[pos] == '\n' || buf[pos] == '\r' || buf[pos] == '\t'))
            ++pos;
        if (b == pos) fail("unexpected end of file");
        return buf.substr(b, pos - b);
    }
    long long ll() { return std::stoll(token()); }
    double dbl() { return std::stod(token()); }
    template <typename T>
    T raw() {
        if (pos + sizeof(T) > buf.size()) fail("unexpected end of binary data");
        T v;
        std::memcpy(&v, buf.data() + pos, sizeof(T));
        pos += sizeof(T);
        return v;
    }
    // size_t-sized integer of MSH 4.1 (data-size 8)
    long long sz() { return binary ? (long long)raw<uint64_t>() : ll(); }
    long long i4() { return binary ? (long long)raw<int32_t>() : ll(); }
    double f8() { return binary ? raw<double>() : dbl(); }
    void expect_end(const std::string& name) {
        // after binary data a newline precedes the end marker
        const std::string marker = "$End" + name;
        skip_ws();
        if (buf.compare(pos, marker.size(), marker) != 0) fail("expected " + marker);
        line();
    }
    // skips a section this reader does not use ($ElementData, $PhysicalNames,
    // ...); after binary data the end marker may follow without a newline
    void skip_section(const std::string& name) {
        const std::string marker = "$End" + name;
        const size_t e = buf.find(marker, pos);
        if (e == std::string::npos) fail("no " + marker);
        pos = e;
        line();
    }
};

void read_format(MshReader& r) {
    std::string h = r.line();
    while (h.empty() && !r.eof()) h = r.line();
    if (h != "$MeshFormat") r.fail("missing $MeshFormat");
    std::istringstream ss(r.line());
    std::string ver;
    int type = 0, dsize = 0;
    ss >> ver >> type >> dsize;
    if (ver.rfind("2.2", 0) == 0)
        r.version = 22;
    else if (ver.rfind("4.1", 0) == 0)
        r.version = 41;
    else
        r.fail("MSH version " + ver + " is not supported (2.2 and 4.1 are)");
    r.binary = type == 1;
    if (r.binary) {
        if (dsize != 8) r.fail("binary MSH with data size " + std::to_string(dsize) + " is not supported");
        if (r.raw<int32_t>() != 1) r.fail("binary MSH of the other byte order is not supported");
        // a newline may or may not follow the binary one (expect_end skips it)
    }
    r.expect_end("MeshFormat");
}

// MSH 4.1 binary $Entities: counts, then per point tag, xyz and physical
// tags, per curve / surface / volume tag, bounding box, physical tags and
// bounding entities (ASCII is skipped as text)
void skip_entities_binary(MshReader& r) {
    const long long np = r.sz(), nc = r.sz(), ns = r.sz(), nvol = r.sz();
    for (long long i = 0; i < np; ++i) {
        r.raw<int32_t>();
        for (int k = 0; k < 3; ++k) r.raw<double>();
        const long long nt = r.sz();
        for (long long t = 0; t < nt; ++t) r.raw<int32_t>();
    }
    for (long long i = 0; i < nc + ns + nvol; ++i) {
        r.raw<int32_t>();
        for (int k = 0; k < 6; ++k) r.raw<double>();
        const long long nt = r.sz();
        for (long long t = 0; t < nt; ++t) r.raw<int32_t>();
        const long long nb = r.sz();
        for (long long t = 0; t < nb; ++t) r.raw<int32_t>();
    }
    r.expect_end("Entities");
}

struct MshData {
    std::vector<double> xyz;                      // node coordinates in file order
    std::unordered_map<long long, int> tag_to_idx;
    std::vector<long long> conn;                  // node tags of the wanted elements
    int node_blocks = 0, elem_blocks = 0;
};

void read_nodes(MshReader& r, MshData& d, bool verbose) {
    if (r.version == 22) {
        const long long n = std::stoll(r.line());
        d.xyz.reserve(3 * n);
        for (long long i = 0; i < n; ++i) {
            long long tag;
            double x, y, z;
            if (r.binary) {
                tag = r.raw<int32_t>();
                x = r.raw<double>();
                y = r.raw<double>();
                z = r.raw<double>();
            } else {
                tag = r.ll();
                x = r.dbl();
                y = r.dbl();
                z = r.dbl();
            }
            d.tag_to_idx[tag] = (int)(d.xyz.size() / 3);
            d.xyz.push_back(x);
            d.xyz.push_back(y);
            d.xyz.push_back(z);
        }
        d.node_blocks = 1;
        r.expect_end("Nodes");
        return;
    }
    const long long nblocks = r.sz(), nnodes = r.sz();
    r.sz();
    r.sz();
    d.xyz.reserve(3 * nnodes);
    for (long long b = 0; b < nblocks; ++b) {
        const long long dim = r.i4();
        r.i4();
        const long long param = r.i4();
        const long long n = r.sz();
        if (param != 0) r.fail("parametric node blocks are not supported");
        if (verbose)
            std::cout << "Node Block " << b << " has " << n << " entries with dimension " << dim << std::endl;
        std::vector<long long> tags((size_t)n);
        for (long long i = 0; i < n; ++i) tags[(size_t)i] = r.sz();
        for (long long i = 0; i < n; ++i) {
            d.tag_to_idx[tags[(size_t)i]] = (int)(d.xyz.size() / 3);
            for (int k = 0; k < 3; ++k) d.xyz.push_back(r.f8());
        }
    }
    d.node_blocks = (int)nblocks;
    r.expect_end("Nodes");
}

void read_elements(MshReader& r, MshData& d, int want, bool verbose) {
    const int npe = gmsh_nodes(want);
    if (r.version == 22) {
        const long long n = std::stoll(r.line());
        long long done = 0;
        while (done < n) {
            if (r.binary) {
                const int type = r.raw<int32_t>();
                const long long cnt = r.raw<int32_t>();
                const int ntags = r.raw<int32_t>();
                const int k = gmsh_nodes(type);
                for (long long e = 0; e < cnt; ++e) {
                    r.raw<int32_t>();
                    for (int t = 0; t < ntags; ++t) r.raw<int32_t>();
                    for (int c = 0; c < k; ++c) {
                        const long long tag = r.raw<int32_t>();
                        if (type == want) d.conn.push_back(tag);
                    }
                }
                done += cnt;
                ++d.elem_blocks;
            } else {
                r.ll();
                const int type = (int)r.ll();
                const int ntags = (int)r.ll();
                for (int t = 0; t < ntags; ++t) r.ll();
                const int k = gmsh_nodes(type);
                for (int c = 0; c < k; ++c) {
                    const long long tag = r.ll();
                    if (type == want) d.conn.push_back(tag);
                }
                ++done;
            }
        }
        if (!r.binary) d.elem_blocks = 1;
        r.expect_end("Elements");
        (void)npe;
        return;
    }
    const long long nblocks = r.sz();
    r.sz();
    r.sz();
    r.sz();
    for (long long b = 0; b < nblocks; ++b) {
        const long long dim = r.i4();
        r.i4();
        const int type = (int)r.i4();
        const long long n = r.sz();
        const int k = gmsh_nodes(type);
        if (verbose) {
            if (type != want)
                std::cout << "Element Block " << b << " is of element type " << type << ", skipping..." << std::endl;
            else
                std::cout << "Element Block " << b << " has " << n << " entries with " << type
                          << " nodes each of dimension " << dim << std::endl;
        }
        for (long long e = 0; e < n; ++e) {
            r.sz();  // element tag
            for (int c = 0; c < k; ++c) {
                const long long tag = r.sz();
                if (type == want) d.conn.push_back(tag);
            }
        }
    }
    d.elem_blocks = (int)nblocks;
    r.expect_end("Elements");
}

bool read_msh(const std::string& filename, int want, Mat& vertices, std::vector<int>& elements, bool verbose) {
    MshReader r;
    r.file = filename;
    {
        std::ifstream f(filename, std::ios::binary);
        if (!f) throw std::runtime_error("cannot open " + filename);
        std::ostringstream ss;
        ss << f.rdbuf();
        r.buf = ss.str();
    }
    read_format(r);
    MshData d;
    bool nodes = false, elems = false;
    while (!r.eof()) {
        const std::string h = r.line();
        if (h.empty()) continue;
        if (h == "$Nodes") {
            read_nodes(r, d, verbose);
            nodes = true;
        } else if (h == "$Elements") {
            if (!nodes) r.fail("$Elements before $Nodes");
            read_elements(r, d, want, verbose);
            elems = true;
        } else if (h == "$Entities" && r.binary) {
            skip_entities_binary(r);
        } else if (h.size() > 1 && h[0] == '$') {
            r.skip_section(h.substr(1));
        } else {
            r.fail("unexpected line: " + h.substr(0, 40));
        }
    }
    if (!nodes || !elems) r.fail("missing $Nodes or $Elements");
    if (verbose && d.node_blocks != 1)
        std::cout << "Multiple node blocks (" << d.node_blocks
                  << ") detected in mesh, only first block of 3D vertices is used..." << std::endl;
    const long nn = (long)(d.xyz.size() / 3);
    vertices = Mat(nn, 3);
    for (long i = 0; i < nn; ++i)
        for (int k = 0; k < 3; ++k) vertices(i, k) = d.xyz[3 * (size_t)i + k];
    elements.resize(d.conn.size());
    for (size_t i = 0; i < d.conn.size(); ++i) {
        auto it = d.tag_to_idx.find(d.conn[i]);
        if (it == d.tag_to_idx.end()) r.fail("element references unknown node tag " + std::to_string(d.conn[i]));
        elements[i] = it->second;
    }
    return true;
}

}  // namespace

bool IO::read_tet_MSH(const std::string& filename, Mat& vertices, std::vector<int>& elements, bool verboseFlag) {
    return read_msh(filename, 4, vertices, elements, verboseFlag);
}

bool IO::read_hex_MSH(const std::string& filename, Mat& vertices, std::vector<int>& elements, bool verboseFlag) {
    return read_msh(filename, 5, vertices, elements, verboseFlag);
}

// ---------------------------------------------------------------- surfaces
namespace {

struct FaceKey {
    std::array<int, 4> v;
    bool operator==(const FaceKey& o) const { return v == o.v; }
};
struct FaceKeyHash {
    size_t operator()(const FaceKey& k) const {
        size_t h = 1469598103934665603ull;
        for (int x : k.v) h = (h ^ (size_t)(uint32_t)x) * 1099511628211ull;
        return h;
    }
};

// io.cpp:570-681 (tet, ed = 4, nf = 3) and 684-780 (hex, ed = 8, nf = 4).
// Boundary faces: every element face in element order, toggled on a set
// keyed by its sorted vertices (a face met twice is removed), the survivors
// in order of first appearance.  Groups: from the first remaining face, the
// earliest remaining face that shares a vertex with the group joins it,
// repeatedly (the reference rescans its list after every addition; a min-heap
// of the faces touching the group gives the same choices).  A face is
// flipped when its normal (v1 - v0) x (v2 - v0) points towards the centroid
// of its element: tet {0, 2, 1}, hex by swapping vertices 1 and 3.
std::tuple<SurfaceGroups, FaceElementGroups> extract_surfaces(const Mat& V, const std::vector<int>& ele, int ed,
                                                              bool verbose) {
    static const int tetF[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
    static const int hexF[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}};
    if (V.cols != 3) throw std::invalid_argument("vertices must be N x 3");
    if (ele.size() % ed != 0) throw std::invalid_argument("eleIdx must be E x " + std::to_string(ed));
    const int ne = (int)(ele.size() / ed), nf = ed == 4 ? 3 : 4, fpe = ed == 4 ? 4 : 6;
    struct Face {
        std::array<int, 4> v;
        V3 com;
        int elem;
        bool alive;
    };
    std::vector<Face> faces;
    std::unordered_map<FaceKey, size_t, FaceKeyHash> live;
    for (int i = 0; i < ne; ++i) {
        const int* e = &ele[(size_t)i * ed];
        for (int j = 0; j < fpe; ++j) {
            std::array<int, 4> f{{-1, -1, -1, -1}};
            for (int k = 0; k < nf; ++k) f[k] = e[ed == 4 ? tetF[j][k] : hexF[j][k]];
            FaceKey key{f};
            std::sort(key.v.begin(), key.v.begin() + nf);
            auto it = live.find(key);
            if (it == live.end()) {
                V3 c{0.0, 0.0, 0.0};
                for (int k = 0; k < ed; ++k)
                    for (int t = 0; t < 3; ++t) c[t] += V(e[k], t);
                for (int t = 0; t < 3; ++t) c[t] /= ed;
                // tets keep the sorted vertex order, hexes the element's face order
                live.emplace(key, faces.size());
                faces.push_back(Face{ed == 4 ? key.v : f, c, i, true});
            } else {
                faces[it->second].alive = false;
                live.erase(it);
            }
        }
    }
    std::vector<size_t> order;
    for (size_t p = 0; p < faces.size(); ++p)
        if (faces[p].alive) order.push_back(p);
    if (verbose)
        std::cout << "Number of surface " << (ed == 4 ? "triangles" : "quads") << ": " << order.size() << std::endl;
    // incidence: vertex -> positions (in the surviving order) of its faces
    const size_t nr = order.size();
    std::unordered_map<int, std::vector<int>> inc;
    for (size_t r = 0; r < nr; ++r)
        for (int k = 0; k < nf; ++k) inc[faces[order[r]].v[k]].push_back((int)r);
    std::vector<char> taken(nr, 0);
    SurfaceGroups groups;
    FaceElementGroups egroups;
    size_t first = 0;
    size_t remaining = nr;
    auto pt = [&](int idx) { return V3{V(idx, 0), V(idx, 1), V(idx, 2)}; };
    while (remaining > 0) {
        while (taken[first]) ++first;
        std::vector<std::vector<int>> g;
        std::vector<int> ge;
        std::set<int> verts;
        std::priority_queue<int, std::vector<int>, std::greater<int>> heap;
        heap.push((int)first);
        while (!heap.empty()) {
            const int r = heap.top();
            heap.pop();
            if (taken[r]) continue;
            taken[r] = 1;
            --remaining;
            const Face& fc = faces[order[r]];
            const V3 a = pt(fc.v[0]), b = pt(fc.v[1]), c = pt(fc.v[2]);
            const V3 n = cross3(sub3(b, a), sub3(c, a));
            std::vector<int> out(fc.v.begin(), fc.v.begin() + nf);
            if (dot3(n, sub3(fc.com, a)) > 0.0) {
                if (ed == 4)
                    out = {fc.v[0], fc.v[2], fc.v[1]};
                else
                    std::swap(out[1], out[3]);
            }
            g.push_back(out);
            ge.push_back(fc.elem);
            for (int k = 0; k < nf; ++k)
                if (verts.insert(fc.v[k]).second)
                    for (int q : inc[fc.v[k]])
                        if (!taken[q]) heap.push(q);
        }
        if (verbose) {
            std::cout << "Group size: " << g.size() << std::endl;
            std::cout << "Remaining faces: " << remaining << std::endl;
        }
        groups.push_back(std::move(g));
        egroups.push_back(std::move(ge));
    }
    return std::make_tuple(groups, egroups);
}

}  // namespace

std::tuple<SurfaceGroups, FaceElementGroups> IO::extract_tet_surfaces(const Mat& vertices,
                                                                      const std::vector<int>& eleIdx,
                                                                      bool verboseFlag) {
    return extract_surfaces(vertices, eleIdx, 4, verboseFlag);
}

std::tuple<SurfaceGroups, FaceElementGroups> IO::extract_hex_surfaces(const Mat& vertices,
                                                                      const std::vector<int>& eleIdx,
                                                                      bool verboseFlag) {
    return extract_surfaces(vertices, eleIdx, 8, verboseFlag);
}

// ---------------------------------------------------------------- grids
// io.cpp:899-928
std::vector<int> IO::hex_to_5_tets(const std::vector<int>& hexes) {
    if (hexes.size() % 8 != 0) throw std::invalid_argument("hexahedra must have 8 vertices");
    static const int pat[5][4] = {{0, 1, 3, 4}, {1, 2, 3, 6}, {1, 4, 5, 6}, {3, 4, 6, 7}, {1, 3, 4, 6}};
    const size_t nh = hexes.size() / 8;
    std::vector<int> t(nh * 20);
    for (size_t j = 0; j < nh; ++j)
        for (int s = 0; s < 5; ++s)
            for (int c = 0; c < 4; ++c) t[20 * j + 4 * s + c] = hexes[8 * j + pat[s][c]];
    return t;
}

// io.cpp:931-987: vid = i + NX (j + NY k)
std::pair<Mat, std::vector<int>> IO::make_voxel_hex_grid(int nx, int ny, int nz, double dx, double dy, double dz) {
    if (nx < 1 || ny < 1 || nz < 1) throw std::runtime_error("nx, ny, nz must be >= 1");
    if (dx <= 0.0 || dy <= 0.0 || dz <= 0.0) throw std::runtime_error("dx, dy, dz must be > 0");
    const int NX = nx + 1, NY = ny + 1, NZ = nz + 1;
    Mat V((long)NX * NY * NZ, 3);
    for (int k = 0; k < NZ; ++k)
        for (int j = 0; j < NY; ++j)
            for (int i = 0; i < NX; ++i) {
                const long id = i + NX * (j + (long)NY * k);
                V(id, 0) = i * dx;
                V(id, 1) = j * dy;
                V(id, 2) = k * dz;
            }
    std::vector<int> H((size_t)nx * ny * nz * 8);
    size_t e = 0;
    auto id = [&](int a, int b, int c) { return a + NX * (b + NY * c); };
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i, ++e) {
                const int c8[8] = {id(i, j, k),         id(i + 1, j, k),         id(i + 1, j + 1, k),
                                   id(i, j + 1, k),     id(i, j, k + 1),         id(i + 1, j, k + 1),
                                   id(i + 1, j + 1, k + 1), id(i, j + 1, k + 1)};
                for (int t = 0; t < 8; ++t) H[8 * e + t] = c8[t];
            }
    return {V, H};
}

}  // namespace sors_host
