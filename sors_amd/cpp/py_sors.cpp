// py_sors.cpp -- pybind11 module with the reference's Python surface
// (cpp/py_sors.cpp:25-261) over the device path: Energy(...), EnergyTet/Hex,
// Solver(settings), SolverTet/Hex.step, Params, SimulationSettings, the IO
// statics (VTU / PVD, Gmsh readers, surfaces, grids), VisualType /
// VisualOption, Plane3D and the DiskContact handle.  numpy arrays are copied in and out like the
// reference's Eigen conversions; step() releases the GIL (an addition).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <array>
#include <cmath>

#include "sors_demos.h"
#include "sors_facade.h"
#include "sors_io.h"

namespace py = pybind11;
using namespace sors_host;

namespace {

Mat to_mat(const py::handle& obj) {
    if (py::isinstance<py::float_>(obj) || py::isinstance<py::int_>(obj)) return Mat(1, 1, obj.cast<double>());
    auto a = py::array_t<double, py::array::forcecast>::ensure(obj);
    if (!a) throw std::invalid_argument("expected a number or an array of numbers");
    if (a.ndim() == 0) return Mat(1, 1, *a.data());
    if (a.ndim() == 1) {
        Mat m((long)a.shape(0), 1);
        auto r = a.unchecked<1>();
        for (long i = 0; i < m.rows; ++i) m(i, 0) = r(i);
        return m;
    }
    if (a.ndim() != 2) throw std::invalid_argument("expected a 1-D or 2-D array");
    Mat m((long)a.shape(0), (long)a.shape(1));
    auto r = a.unchecked<2>();
    for (long i = 0; i < m.rows; ++i)
        for (long j = 0; j < m.cols; ++j) m(i, j) = r(i, j);
    return m;
}

py::array_t<double> from_mat(const Mat& m) {
    py::array_t<double> out({(py::ssize_t)m.rows, (py::ssize_t)m.cols});
    auto w = out.mutable_unchecked<2>();
    for (long i = 0; i < m.rows; ++i)
        for (long j = 0; j < m.cols; ++j) w(i, j) = m(i, j);
    return out;
}

std::vector<double> to_vec(const py::handle& obj) {
    auto a = py::array_t<double, py::array::forcecast>::ensure(obj);
    if (!a) throw std::invalid_argument("expected an array of numbers");
    std::vector<double> v((size_t)a.size());
    const double* p = a.data();
    // element order of a C-contiguous copy (row-major flattening)
    auto c = py::array_t<double, py::array::c_style | py::array::forcecast>::ensure(obj);
    p = c.data();
    for (size_t i = 0; i < v.size(); ++i) v[i] = p[i];
    return v;
}

py::array_t<double> from_vec(const std::vector<double>& v) {
    py::array_t<double> out((py::ssize_t)v.size());
    std::copy(v.begin(), v.end(), out.mutable_data());
    return out;
}

std::vector<int> conn_of(const py::handle& obj, int E) {
    auto a = py::array_t<int, py::array::c_style | py::array::forcecast>::ensure(obj);
    if (!a || a.ndim() != 2 || a.shape(1) != E)
        throw std::invalid_argument("eleIdx must be a 2-D integer array with " + std::to_string(E) + " columns");
    return std::vector<int>(a.data(), a.data() + a.size());
}

template <int E>
Energy<E>* make_energy(const py::handle& V, const py::handle& eleIdx, const SimulationSettings& settings,
                       const py::handle& grav, double alpha, const std::vector<std::set<std::string>>& ee,
                       const std::vector<Params>& ep, const std::vector<std::string>& ct, const Params& cp,
                       const std::vector<std::string>& ft, const Params& fp) {
    Mat X = to_mat(V);
    auto g = to_vec(grav);
    if (g.size() != 3) throw std::invalid_argument("gravAcceleration must have 3 entries");
    double gv[3] = {g[0], g[1], g[2]};
    return new Energy<E>(X, conn_of(eleIdx, E), settings, gv, alpha, ee, ep, ct, cp, ft, fp);
}

template <int E>
void bind_energy(py::module_& m, const char* name) {
    py::class_<Energy<E>>(m, name)
        .def(py::init(&make_energy<E>), py::arg("undeformedVertices"), py::arg("eleIdx"), py::arg("settings"),
             py::arg("gravAcceleration"), py::arg("dampingAlpha") = 0.0, py::arg("elementEnergiesList"),
             py::arg("elementParameterList"), py::arg("constraintTypesList"), py::arg("constraintParameterList"),
             py::arg("forceTypesList"), py::arg("forceParameterList"))
        .def("set_initial_velocity", [](Energy<E>& e, const py::handle& v) { e.set_initial_velocity(to_mat(v)); })
        .def("set_initial_deformation",
             [](Energy<E>& e, const py::handle& v) { e.set_initial_deformation(to_mat(v)); })
        .def("get_vPrev", [](const Energy<E>& e) { return from_vec(e.get_vPrev()); })
        .def("compute_energy",
             [](Energy<E>& e, const py::handle& q, double dt, const Params& act) {
                 return e.compute_energy(to_vec(q), dt, act);
             },
             py::arg("x"), py::arg("dt"), py::arg("actuation") = Params())
        .def("compute_gradient",
             [](Energy<E>& e, const py::handle& q, double dt, const Params& act) {
                 return from_vec(e.compute_gradient(to_vec(q), dt, act));
             },
             py::arg("x"), py::arg("dt"), py::arg("actuation") = Params());
}

template <int E>
void bind_solver(py::module_& m, const char* name) {
    py::class_<Solver<E>>(m, name)
        .def(
            "step",
            [](const Solver<E>& s, const py::handle& sol, Energy<E>& en, const Params& act, double dt, int sub) {
                auto q = to_vec(sol);
                std::vector<double> out;
                {
                    py::gil_scoped_release nogil;
                    out = s.step(q, en, act, dt, sub);
                }
                return from_vec(out);
            },
            "Step solution forward in time.", py::arg("solution"), py::arg("systemEnergy"), py::arg("act"),
            py::arg("dt") = 0.0, py::arg("substeps") = 0)
        .def(
            "step",
            [](const Solver<E>& s, const py::handle& sol, Energy<E>& en, double dt, int sub) {
                auto q = to_vec(sol);
                std::vector<double> out;
                {
                    py::gil_scoped_release nogil;
                    out = s.step(q, en, dt, sub);
                }
                return from_vec(out);
            },
            "Step solution forward in time with default actuation.", py::arg("solution"), py::arg("systemEnergy"),
            py::arg("dt") = 0.0, py::arg("substeps") = 0)
        .def_property_readonly("last_stats", [](const Solver<E>& s) {
            const sors_stats& t = s.last_stats;
            py::dict d;
            d["newton_iters"] = t.newton_iters;
            d["pcg_iters"] = t.pcg_iters;
            d["backtracks"] = t.backtracks;
            d["gnorm"] = t.gnorm;
            d["dxnorm"] = t.dxnorm;
            d["ms_total"] = t.ms_total;
            d["fp64_fallbacks"] = t.fp64_fallbacks;
            return d;
        });
}

std::vector<int> int_rows(const py::handle& obj, int cols, const char* what) {
    auto a = py::array_t<int, py::array::c_style | py::array::forcecast>::ensure(obj);
    if (!a || a.ndim() != 2 || a.shape(1) != cols)
        throw std::invalid_argument(std::string(what) + " must be a 2-D integer array with " + std::to_string(cols) +
                                    " columns");
    return std::vector<int>(a.data(), a.data() + a.size());
}

py::array_t<int> int_array(const std::vector<int>& v, int cols) {
    py::array_t<int> out({(py::ssize_t)(v.size() / cols), (py::ssize_t)cols});
    std::copy(v.begin(), v.end(), out.mutable_data());
    return out;
}

std::array<double, 3> vec3(const py::handle& v) {
    auto a = to_vec(v);
    if (a.size() != 3) throw std::invalid_argument("expected a 3-vector");
    return {a[0], a[1], a[2]};
}

py::array_t<double> arr3(const std::array<double, 3>& a) { return from_vec({a[0], a[1], a[2]}); }

template <int E>
DiskContactHandle<E> disk_handle(Energy<E>& en, int idx) {
    if (idx < 0 || idx >= (int)en.constraintTypesList_.size()) throw py::index_error("constraint index out of range");
    if (en.constraintTypesList_[idx] != "diskContact") throw py::value_error("constraint at index is not DiskContact");
    return DiskContactHandle<E>{&en};
}

template <int E>
void bind_disk(py::module_& m, const char* name) {
    py::class_<DiskContactHandle<E>>(m, name)
        .def(
            "set_disk",
            [](DiskContactHandle<E>& h, const py::handle& c, const py::handle& nrm, double r, int idx) {
                auto cc = vec3(c), nn = vec3(nrm);
                h.set_disk(cc.data(), nn.data(), r, idx);
            },
            py::arg("center"), py::arg("normal"), py::arg("radius"), py::arg("diskIdx") = 0);
    m.def(
        "get_disk_contact_constraint", [](Energy<E>& en, int idx) { return disk_handle<E>(en, idx); },
        py::arg("energy"), py::arg("idx"), py::keep_alive<0, 1>());
}

}  // namespace

PYBIND11_MODULE(py_sors, m) {
    m.doc() = "SORS implicit Newton step on MI355X (HIP device path behind the reference's interface)";
    // py_sors.cpp:27
    m.def("run_msd_timestepping", &run_msd_timestepping, "Run msd demo", py::arg("meshFilePath"),
          py::arg("outputFolder"), py::arg("numSteps"), py::call_guard<py::gil_scoped_release>());

    py::class_<Params>(m, "Params")
        .def(py::init<>())
        .def(py::init([](const std::vector<std::string>& keys, const py::list& values) {
            if (keys.size() != values.size()) throw std::invalid_argument("keys and values differ in length");
            Params p;
            for (size_t i = 0; i < keys.size(); ++i) p.set_param(keys[i], to_mat(values[i]));
            return p;
        }))
        .def("get_value", [](const Params& p, const std::string& k) { return from_mat(p.get_value(k)); })
        .def("set_param", [](Params& p, const std::string& k, const py::object& v) { p.set_param(k, to_mat(v)); })
        .def("set_param", [](Params& p, const std::string& filename) { p.set_param(filename); });

    py::class_<SimulationSettings>(m, "SimulationSettings")
        .def(py::init<>())
        .def_readwrite("description", &SimulationSettings::description)
        .def_readwrite("verbose", &SimulationSettings::verbose)
        .def_readwrite("meshFilePath", &SimulationSettings::meshFilePath)
        .def_readwrite("outputFolder", &SimulationSettings::outputFolder)
        .def_readwrite("meshFileName", &SimulationSettings::meshFileName)
        .def_readwrite("meshType", &SimulationSettings::meshType)
        .def_readwrite("dt", &SimulationSettings::dt)
        .def_readwrite("numSteps", &SimulationSettings::numSteps)
        .def_readwrite("substeps", &SimulationSettings::substeps)
        .def_readwrite("CFLtimeSteppingFlag", &SimulationSettings::CFLtimeSteppingFlag)
        .def_readwrite("timeSteppingScheme", &SimulationSettings::timeSteppingScheme)
        .def_readwrite("solverMethod", &SimulationSettings::solverMethod)
        .def_readwrite("numThreads", &SimulationSettings::numThreads)
        .def_readwrite("precision", &SimulationSettings::precision)
        .def_readwrite("pcgTolerance", &SimulationSettings::pcgTolerance)
        .def_readwrite("predictor", &SimulationSettings::predictor)
        .def_readwrite("coarseNodes", &SimulationSettings::coarseNodes)
        .def_readwrite("coarseEvery", &SimulationSettings::coarseEvery)
        .def_readwrite("contactModel", &SimulationSettings::contactModel)
        .def("verbose_print", &SimulationSettings::verbose_print);

    bind_energy<4>(m, "EnergyTet");
    bind_energy<8>(m, "EnergyHex");
    bind_solver<4>(m, "SolverTet");
    bind_solver<8>(m, "SolverHex");

    // py_sors.cpp:30-62: dispatch on settings.meshType
    m.def(
        "Energy",
        [](const py::handle& V, const py::handle& eleIdx, const SimulationSettings& settings, const py::handle& grav,
           double alpha, const std::vector<std::set<std::string>>& ee, const std::vector<Params>& ep,
           const std::vector<std::string>& ct, const Params& cp, const std::vector<std::string>& ft,
           const Params& fp) -> py::object {
            if (settings.meshType == "Tetrahedron")
                return py::cast(make_energy<4>(V, eleIdx, settings, grav, alpha, ee, ep, ct, cp, ft, fp),
                                py::return_value_policy::take_ownership);
            if (settings.meshType == "Hexahedron")
                return py::cast(make_energy<8>(V, eleIdx, settings, grav, alpha, ee, ep, ct, cp, ft, fp),
                                py::return_value_policy::take_ownership);
            throw std::runtime_error("Unknown mesh type: " + settings.meshType);
        });
    // py_sors.cpp:99-111
    m.def("Solver", [](const SimulationSettings& settings) -> py::object {
        if (settings.meshType == "Tetrahedron")
            return py::cast(new Solver<4>(settings), py::return_value_policy::take_ownership);
        if (settings.meshType == "Hexahedron")
            return py::cast(new Solver<8>(settings), py::return_value_policy::take_ownership);
        throw std::runtime_error("Unknown mesh type: " + settings.meshType);
    });

    // py_sors.cpp:201-227
    py::enum_<VisualType>(m, "VisualType")
        .value("VectorVertex", VisualType::VectorVertex)
        .value("VectorElement", VisualType::VectorElement)
        .value("ScalarVertex", VisualType::ScalarVertex)
        .value("ScalarElement", VisualType::ScalarElement);
    py::class_<VisualOption>(m, "VisualOption")
        .def(py::init<>())
        .def_readwrite("name", &VisualOption::name)
        .def_readwrite("type", &VisualOption::type)
        .def_property(
            "data", [](const VisualOption& v) { return from_mat(v.data); },
            [](VisualOption& v, const py::handle& d) { v.data = to_mat(d); })
        .def_readwrite("threshold", &VisualOption::threshold);
    py::class_<Plane3D>(m, "Plane3D")
        .def(py::init<>())
        .def(py::init([](const py::handle& point, const py::handle& normal) {
                 return Plane3D(vec3(point), vec3(normal));
             }),
             py::arg("point"), py::arg("normal"))
        .def("is_point_on_plane", [](const Plane3D& pl, const py::handle& x) { return pl.is_point_on_plane(vec3(x)); })
        .def("distance_to_point", [](const Plane3D& pl, const py::handle& x) { return pl.distance_to_point(vec3(x)); })
        .def("project_point_on_plane",
             [](const Plane3D& pl, const py::handle& x) { return arr3(pl.project_point_on_plane(vec3(x))); })
        .def("create_square_points", &Plane3D::create_square_points, py::arg("squareSide"))
        .def("get_point", [](const Plane3D& pl) { return arr3(pl.get_point()); })
        .def("set_point", [](Plane3D& pl, const py::handle& p) { pl.set_point(vec3(p)); })
        .def("get_normal", [](const Plane3D& pl) { return arr3(pl.get_normal()); })
        .def("set_normal", [](Plane3D& pl, const py::handle& n) { pl.set_normal(vec3(n)); })
        .def("get_square_points", &Plane3D::get_square_points);

    bind_disk<4>(m, "DiskContact");
    bind_disk<8>(m, "DiskContactHex");

    // py_sors.cpp:180-198 (read_*_MSH and make_voxel_hex_grid are additions:
    // the reference exposes them to C++ only)
    py::class_<IO>(m, "IO")
        .def_static(
            "save_tet_VTU",
            [](const std::string& f, const py::handle& V, const py::handle& T, const std::vector<VisualOption>& vis,
               const std::vector<Plane3D>& planes) {
                return IO::save_tet_VTU(f, to_mat(V), int_rows(T, 4, "elements"), vis, planes);
            },
            "Save tetrahedral mesh (vertices and elements) to VTU file.", py::arg("filename"), py::arg("vertices"),
            py::arg("elements"), py::arg("visuals") = std::vector<VisualOption>(),
            py::arg("planes") = std::vector<Plane3D>())
        .def_static(
            "save_hex_VTU",
            [](const std::string& f, const py::handle& V, const py::handle& H, const std::vector<VisualOption>& vis,
               const std::vector<Plane3D>& planes) {
                return IO::save_hex_VTU(f, to_mat(V), int_rows(H, 8, "elements"), vis, planes);
            },
            "Save hexahedral mesh (vertices and elements) to VTU file.", py::arg("filename"), py::arg("vertices"),
            py::arg("elements"), py::arg("visuals") = std::vector<VisualOption>(),
            py::arg("planes") = std::vector<Plane3D>())
        .def_static("save_pvd", &IO::save_pvd, "Save PVD file summarizing all VTUs.", py::arg("filename"),
                    py::arg("outputFolder"), py::arg("timesteps") = 1, py::arg("dt") = 1.0)
        .def_static(
            "extract_tet_surfaces",
            [](const py::handle& V, const py::handle& T, bool verbose) {
                return IO::extract_tet_surfaces(to_mat(V), int_rows(T, 4, "eleIdx"), verbose);
            },
            "Extract surfaces from tetrahedral mesh.", py::arg("vertices"), py::arg("eleIdx"),
            py::arg("verboseFlag") = false)
        .def_static(
            "extract_hex_surfaces",
            [](const py::handle& V, const py::handle& H, bool verbose) {
                return IO::extract_hex_surfaces(to_mat(V), int_rows(H, 8, "eleIdx"), verbose);
            },
            "Extract surfaces from hexahedral mesh.", py::arg("vertices"), py::arg("eleIdx"),
            py::arg("verboseFlag") = false)
        .def_static(
            "read_tet_MSH",
            [](const std::string& f, bool verbose) {
                Mat V;
                std::vector<int> T;
                IO::read_tet_MSH(f, V, T, verbose);
                return py::make_tuple(from_mat(V), int_array(T, 4));
            },
            "Read the vertices and tetrahedra of a Gmsh .msh file (2.2 / 4.1, ASCII / binary).", py::arg("filename"),
            py::arg("verboseFlag") = false)
        .def_static(
            "read_hex_MSH",
            [](const std::string& f, bool verbose) {
                Mat V;
                std::vector<int> H;
                IO::read_hex_MSH(f, V, H, verbose);
                return py::make_tuple(from_mat(V), int_array(H, 8));
            },
            "Read the vertices and hexahedra of a Gmsh .msh file (2.2 / 4.1, ASCII / binary).", py::arg("filename"),
            py::arg("verboseFlag") = false)
        .def_static(
            "make_voxel_hex_grid",
            [](int nx, int ny, int nz, double dx, double dy, double dz) {
                auto g = IO::make_voxel_hex_grid(nx, ny, nz, dx, dy, dz);
                return py::make_tuple(from_mat(g.first), int_array(g.second, 8));
            },
            py::arg("nx"), py::arg("ny"), py::arg("nz"), py::arg("dx"), py::arg("dy"), py::arg("dz"))
        .def_static(
            "hex_to_5_tets",
            [](const py::handle& H) { return int_array(IO::hex_to_5_tets(int_rows(H, 8, "hexahedra")), 4); },
            "Converts hexahedral elements to 5 tetrahedra.", py::arg("hexElements"));
}
