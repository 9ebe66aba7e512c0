// sors_io.h -- the reference's IO statics and visualisation types, host side
// (cpp/headers/io.h:25-270, cpp/src/io.cpp:3-780, cpp/headers/common.h:53-63,
// cpp/headers/geometries/plane3D.h:15-89), without MshIO or vtu11:
//   - Gmsh .msh readers, formats 2.2 and 4.1, ASCII and binary;
//   - VTU (ASCII XML UnstructuredGrid) and PVD writers with the reference's
//     datasets, thresholds and plane quads;
//   - boundary-surface extraction with the reference's face order, grouping
//     and outward orientation;
//   - the structured-grid helpers (make_voxel_hex_grid, hex_to_5_tets).
// Connectivity arrays are row-major int vectors of E x ed.
#pragma once

#include <array>
#include <string>
#include <tuple>
#include <vector>

#include "sors_facade.h"

namespace sors_host {

// common.h:53-63
enum class VisualType { VectorVertex, VectorElement, ScalarVertex, ScalarElement };
struct VisualOption {
    std::string name;
    VisualType type = VisualType::ScalarVertex;
    Mat data;  // 1 x N / 1 x E (scalar) or 3 x N / 3 x E (vector)
    double threshold = -std::numeric_limits<double>::infinity();
};

// plane3D.h:15-89: a point and a unit normal, and a 1 m square on the plane
// around the point for visualisation
class Plane3D {
public:
    Plane3D();
    Plane3D(const std::array<double, 3>& point, const std::array<double, 3>& normal);
    bool is_point_on_plane(const std::array<double, 3>& x) const;
    double distance_to_point(const std::array<double, 3>& x) const;
    std::array<double, 3> project_point_on_plane(const std::array<double, 3>& x) const;
    std::vector<double> create_square_points(double squareSide) const;
    std::array<double, 3> get_point() const { return point_; }
    void set_point(const std::array<double, 3>& p);
    std::array<double, 3> get_normal() const { return normal_; }
    void set_normal(const std::array<double, 3>& n);
    std::vector<double> get_square_points() const { return squarePoints_; }

private:
    std::array<double, 3> point_{{0.0, 0.0, 0.0}};
    std::array<double, 3> normal_{{0.0, 0.0, 1.0}};
    std::vector<double> squarePoints_;
};

using SurfaceGroups = std::vector<std::vector<std::vector<int>>>;
using FaceElementGroups = std::vector<std::vector<int>>;

struct IO {
    // io.cpp:3-175 / 178-336: vertices N x 3, elements row-major E x 4 / E x 8
    static bool save_tet_VTU(const std::string& filename, const Mat& vertices, const std::vector<int>& elements,
                             std::vector<VisualOption> visuals = {}, const std::vector<Plane3D>& planes = {});
    static bool save_hex_VTU(const std::string& filename, const Mat& vertices, const std::vector<int>& elements,
                             std::vector<VisualOption> visuals = {}, const std::vector<Plane3D>& planes = {});
    // io.cpp:339-356
    static bool save_pvd(const std::string& filename, const std::string& outputFolder, int timesteps = 1,
                         double dt = 1.0);
    // io.cpp:359-568: every node block in order; the tetrahedra (Gmsh type 4)
    // or hexahedra (type 5) of every element block in order, node tags mapped
    // to vertex indices.  Throws std::runtime_error on malformed input.
    static bool read_tet_MSH(const std::string& filename, Mat& vertices, std::vector<int>& elements,
                             bool verboseFlag = false);
    static bool read_hex_MSH(const std::string& filename, Mat& vertices, std::vector<int>& elements,
                             bool verboseFlag = false);
    // io.cpp:570-681 / 684-780: boundary faces grouped into connected
    // surfaces, each face oriented outward, with the element of each face
    static std::tuple<SurfaceGroups, FaceElementGroups> extract_tet_surfaces(const Mat& vertices,
                                                                             const std::vector<int>& eleIdx,
                                                                             bool verboseFlag = false);
    static std::tuple<SurfaceGroups, FaceElementGroups> extract_hex_surfaces(const Mat& vertices,
                                                                             const std::vector<int>& eleIdx,
                                                                             bool verboseFlag = false);
    // io.cpp:899-928
    static std::vector<int> hex_to_5_tets(const std::vector<int>& hexElements);
    // io.cpp:931-987: vertices ((nx+1)(ny+1)(nz+1) x 3) and hexes (nx ny nz x 8)
    static std::pair<Mat, std::vector<int>> make_voxel_hex_grid(int nx, int ny, int nz, double dx, double dy,
                                                                double dz);
};

}  // namespace sors_host
