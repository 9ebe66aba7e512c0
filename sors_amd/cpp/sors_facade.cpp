// sors_facade.cpp -- host interfaces of the reference over the device C ABI.
#include "sors_facade.h"

#include <algorithm>

#include "sors_elements.h"

#include <cmath>

#include <fstream>
#include <iomanip>
#include <sstream>

namespace sors_host {

// ---------------------------------------------------------------- Params
Params::Params(const std::vector<std::string>& keys, const std::vector<Mat>& values) {
    if (keys.size() != values.size()) throw std::invalid_argument("Params: keys and values differ in length");
    for (size_t i = 0; i < keys.size(); ++i) set_param(keys[i], values[i]);
}

Params::Params(const std::vector<std::string>& keys, const std::vector<double>& values) {
    if (keys.size() != values.size()) throw std::invalid_argument("Params: keys and values differ in length");
    for (size_t i = 0; i < keys.size(); ++i) set_param(keys[i], values[i]);
}

Mat Params::get_value(const std::string& key) const {
    auto it = dict_.find(key);
    if (it == dict_.end()) throw std::out_of_range("Key not found in the parameter list: " + key);
    return it->second;
}

// params.h:43-49: an existing key keeps its size
void Params::set_param(const std::string& key, const Mat& value) {
    auto it = dict_.find(key);
    if (it != dict_.end() && it->second.size() != value.size())
        throw std::invalid_argument("Params: new value for '" + key + "' changes its size");
    dict_[key] = value;
}

void Params::set_param(const std::string& key, const std::vector<double>& values) {
    Mat m((long)values.size(), 1);
    for (size_t i = 0; i < values.size(); ++i) m.data[i] = values[i];
    set_param(key, m);
}

void Params::set_param(const std::string& key, double value) {
    Mat m(1, 1, value);
    set_param(key, m);
}

// params.h:118-174: CSV lines "key,v1[,v2...]"
void Params::set_param(const std::string& filename) {
    if (filename.size() < 4 || filename.substr(filename.find_last_of('.') + 1) != "csv")
        throw std::invalid_argument("Could not load this type of file: " + filename);
    std::ifstream f(filename);
    if (!f.is_open()) throw std::invalid_argument("Could not open the file: " + filename);
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty()) continue;
        std::vector<std::string> cells;
        std::stringstream ss(line);
        std::string cell;
        while (std::getline(ss, cell, ',')) cells.push_back(cell);
        if (cells.size() < 2) throw std::invalid_argument("Params CSV line needs a key and a value: " + line);
        if (cells.size() > 2) {
            std::vector<double> v;
            for (size_t i = 1; i < cells.size(); ++i) v.push_back(std::stod(cells[i]));
            set_param(cells[0], v);
        } else {
            set_param(cells[0], std::stod(cells[1]));
        }
    }
}

// ---------------------------------------------------------------- settings
void SimulationSettings::verbose_print() const {
    auto na = [](const std::string& s) { return s.empty() ? std::string("not indicated") : s; };
    std::cout << "\n============================ Simulation Description ============================\n"
              << na(description) << "\n"
              << "\n============================== Simulation Settings =============================\n";
    std::cout << std::left << std::setw(25) << "Mesh File Path:" << na(meshFilePath) << "\n"
              << std::setw(25) << "Output Folder:" << na(outputFolder) << "\n"
              << std::setw(25) << "Mesh File Name:" << na(meshFileName) << "\n"
              << std::setw(25) << "Mesh Type" << na(meshType) << "\n"
              << std::setw(25) << "Time Step (dt):";
    if (std::isnan(dt))
        std::cout << "not indicated\n";
    else
        std::cout << dt << "\n";
    std::cout << std::setw(25) << "Number of Steps:" << (numSteps < 0 ? "not indicated" : std::to_string(numSteps))
              << "\n"
              << std::setw(25) << "Number of Substeps:" << substeps << "\n"
              << std::setw(25) << "CFL Timestepping:" << (CFLtimeSteppingFlag ? "On" : "Off") << "\n"
              << std::setw(25) << "Time Scheme:" << na(timeSteppingScheme) << "\n"
              << std::setw(25) << "Solver Method:" << na(solverMethod) << "\n"
              << std::setw(25) << "Precision:" << precision << "\n"
              << std::setw(25) << "Contact Model:" << contactModel << "\n"
              << std::setw(25) << "PCG Coarse Space:"
              << (coarseNodes > 0 ? std::to_string(coarseNodes) + " nodes per aggregate" : std::string("off")) << "\n"
              << "================================================================================\n";
}

// ---------------------------------------------------------------- errors
void check_status(int code, const sors_ctx* ctx) {
    if (code == SORS_OK) return;
    std::string detail = sors_last_error(ctx);
    switch (code) {
        case SORS_ERR_INVERTED: throw std::runtime_error("Inverted Elements");
        case SORS_ERR_NOCONV: throw std::runtime_error("Newton's method did not converge");
        case SORS_ERR_NAN: throw std::runtime_error("Newton's method did not converge (" + detail + ")");
        case SORS_ERR_LINESEARCH: throw std::runtime_error("Newton's method did not converge (" + detail + ")");
        case SORS_ERR_UNSUPPORTED: throw std::runtime_error("Solver method not recognized (" + detail + ")");
        default: throw std::runtime_error(detail.empty() ? "sors device error" : detail);
    }
}

// ---------------------------------------------------------------- Energy
template <int E>
Energy<E>::Energy(const Mat& X, const std::vector<int>& conn, const SimulationSettings& settings,
                  const double grav[3], double dampingAlpha,
                  const std::vector<std::set<std::string>>& elementEnergiesList,
                  const std::vector<Params>& elementParameterList, const std::vector<std::string>& constraintTypesList,
                  const Params& constraintParameterList, const std::vector<std::string>& forceTypesList,
                  const Params& forceParameterList)
    : settings_(settings), forceTypesList_(forceTypesList) {
    if (X.cols != 3) throw std::invalid_argument("undeformedVertices must be N x 3");
    numVertices_ = (int)X.rows;
    const int nel = (int)(conn.size() / E);
    if ((long)conn.size() != (long)nel * E) throw std::invalid_argument("eleIdx size is not a multiple of E");
    if ((int)elementEnergiesList.size() != nel || (int)elementParameterList.size() != nel)
        throw std::invalid_argument("one energy set and one Params per element are required");
    std::vector<double> rest = X.row_major();
    std::vector<double> Y(nel), P(nel), D(nel);
    // every element's energy plugins, built by name through the registry
    // (tetrahedron.cpp:55-95, hexahedron.cpp:24-64); their descriptions give
    // the device arrays.  One elastic energy per element ("neohookean" or
    // "stableneohookean", the same type on every element), plus
    // "softconMuscle" on any subset of elements
    int material = -1;
    std::vector<int> mgroup(nel, -1);
    std::vector<double> mk(nel, 0.0), mdir(3 * (size_t)nel, 0.0);
    int mgroups = 0;
    elements_.reserve((size_t)nel);
    for (int e = 0; e < nel; ++e) {
        const auto& set = elementEnergiesList[e];
        if (set.empty()) throw std::runtime_error("element " + std::to_string(e) + " has no energy");
        Mat Xe(E, 3);
        for (int a = 0; a < E; ++a) {
            const int v = conn[(size_t)E * e + a];
            if (v < 0 || v >= numVertices_) throw std::invalid_argument("eleIdx refers to a vertex out of range");
            for (int k = 0; k < 3; ++k) Xe(a, k) = X(v, k);
        }
        elements_.push_back(std::make_unique<ElementType<E>>(elementParameterList[e], Xe, set));
        const DeviceElementDesc dd = elements_.back()->device_description();
        if (dd.material < 0)
            throw std::runtime_error("element " + std::to_string(e) +
                                     ": the device path needs an elastic energy (neohookean or stableneohookean)");
        if (material >= 0 && dd.material != material)
            throw std::runtime_error("the device path takes one elastic energy type for all elements");
        material = dd.material;
        if (dd.muscleGroup >= 0) {
            mgroup[e] = dd.muscleGroup;
            mk[e] = dd.muscleStiffness;
            for (int k = 0; k < 3; ++k) mdir[3 * (size_t)e + k] = dd.muscleDirection[k];
            mgroups = std::max(mgroups, dd.muscleGroup + 1);
        }
        D[e] = elements_.back()->get_density();
        Y[e] = dd.youngs;
        P[e] = dd.poisson;
    }
    if (material < 0) material = SORS_NEOHOOKEAN;
    std::vector<int> mask;
    std::vector<double> bcv;
    // planeContact / diskContact (energy.cpp:69-105): constraints of
    // minimize_SQP, or (contactModel "penalty") a penalty + lagged Coulomb
    // friction (sors_set_contact)
    std::vector<double> pp, pn, dc, dn, dr;
    std::vector<int> surf;
    bool haveSurf = false;
    constraintTypesList_ = constraintTypesList;
    auto flat = [&](const char* key) { return constraintParameterList.get_value(key).row_major(); };
    for (const auto& c : constraintTypesList) {
        if (c == "planeContact" || c == "diskContact") {
            if (c == "planeContact") {
                pp = flat("positionVectorListPlanes");
                pn = flat("normalVectorListPlanes");
                if (pp.size() != pn.size() || pp.size() % 3 != 0)
                    throw std::invalid_argument("planeContact: positions / normals must be lists of 3-vectors");
            } else {
                dc = flat("diskCenterVectorList");
                dn = flat("diskNormalVectorList");
                dr = flat("diskRadiusList");
                if (dc.size() != dn.size() || dc.size() % 3 != 0 || dr.size() * 3 != dc.size())
                    throw std::invalid_argument("DiskContact: centers/normals must be same length and multiples of 3.");
            }
            auto sv = flat("surfaceVerticesIdx");
            if (sv.empty()) throw std::invalid_argument("Surface vertices indices cannot be empty for contact.");
            if (!haveSurf) {
                for (double v : sv) surf.push_back((int)v);
                haveSurf = true;
            }
            hasContact_ = true;
            continue;
        }
        if (c == "neumannBC") {
            // energy.cpp:61-68
            Mat m = constraintParameterList.get_value("neumannBCmask");
            Mat v = constraintParameterList.get_value("neumannBCvalue");
            auto mr = m.row_major(), vr = v.row_major();
            if ((int)mr.size() != 3 * numVertices_ || (int)vr.size() != 3 * numVertices_)
                throw std::invalid_argument("neumannBCmask / neumannBCvalue must have 3N entries");
            mask.resize(mr.size());
            for (size_t i = 0; i < mr.size(); ++i) mask[i] = (int)mr[i];
            bcv = vr;
        } else {
            throw std::runtime_error("Unknown constraint type: " + c +
                                     " (the device path implements neumannBC, planeContact, diskContact)");
        }
    }
    // external forces (energy.cpp:155-172): "vertexForce" (the actuation
    // vector) and "pressure" (PressureForce on "surfaceVertexIdx" faces in
    // "surfaceGroups")
    std::vector<int> pfaces, pgroups;
    for (const auto& f : forceTypesList) {
        if (f == "vertexForce") continue;
        if (f == "pressure") {
            const Mat fv = forceParameterList.get_value("surfaceVertexIdx");
            const std::vector<double> g = forceParameterList.get_value("surfaceGroups").row_major();  // reshaped<RowMajor>()
            const long nv = E == 8 ? 4 : 3;
            if (fv.cols != nv) throw std::invalid_argument("surfaceVertexIdx must have " + std::to_string(nv) + " columns");
            if ((long)g.size() != fv.rows) throw std::invalid_argument("one surfaceGroups entry per face is required");
            for (double v : fv.row_major()) pfaces.push_back((int)v);
            for (double v : g) pgroups.push_back((int)v);
            pressureGroups_ = 0;
            for (int v : pgroups) pressureGroups_ = std::max(pressureGroups_, v + 1);
            continue;
        }
        throw std::runtime_error("Unknown force type: " + f + " (the device path implements vertexForce, pressure)");
    }
    sors_system_desc d{};
    d.n_nodes = numVertices_;
    d.n_elems = nel;
    d.nodes_per_elem = E;
    d.rest = rest.data();
    d.conn = conn.data();
    d.youngs = Y.data();
    d.poisson = P.data();
    d.density = D.data();
    d.material = material;
    d.bc_mask = mask.empty() ? nullptr : mask.data();
    d.bc_value = bcv.empty() ? nullptr : bcv.data();
    for (int k = 0; k < 3; ++k) d.grav[k] = grav[k];
    d.damping_alpha = dampingAlpha;
    if (settings.contactModel != "constraints" && settings.contactModel != "penalty")
        throw std::invalid_argument("contactModel must be \"constraints\" or \"penalty\"");
    const bool penaltyContact = hasContact_ && settings.contactModel == "penalty";
    if (settings.timeSteppingScheme == "backward_euler")
        d.scheme = SORS_BACKWARD_EULER;
    else if (settings.timeSteppingScheme == "crank_nicolson")
        d.scheme = SORS_CRANK_NICOLSON;
    else
        throw std::runtime_error("Unknown time stepping scheme: " + settings.timeSteppingScheme);
    if (settings.solverMethod == "minimize_newton")
        d.method = SORS_MINIMIZE_NEWTON;
    else if (settings.solverMethod == "minimize_newton_linesearch")
        d.method = SORS_MINIMIZE_NEWTON_LINESEARCH;
    else if (settings.solverMethod == "minimize_frozenhessian")
        d.method = SORS_MINIMIZE_FROZENHESSIAN;
    else if (settings.solverMethod == "minimize_SQP")
        // the reference's contact solver.  With contactModel "penalty" and
        // contact the device runs the penalty formulation under Newton + line
        // search, or plain Newton when a muscle is present (its energy is not
        // the potential of its gradient, softconMuscle.cpp:31 vs :72, so an
        // Armijo test on it is meaningless; SQP never evaluates the energy)
        d.method = penaltyContact ? (mgroups > 0 ? SORS_MINIMIZE_NEWTON : SORS_MINIMIZE_NEWTON_LINESEARCH)
                                  : SORS_MINIMIZE_SQP;
    else
        throw std::runtime_error("Solver method not recognized");
    sqpMethod_ = d.method == SORS_MINIMIZE_SQP ? SORS_MINIMIZE_SQP
                                               : (mgroups > 0 ? SORS_MINIMIZE_NEWTON : SORS_MINIMIZE_NEWTON_LINESEARCH);
    check_status(sors_create(&d, &ctx_), nullptr);
    if (mgroups > 0) {
        sors_muscle_desc md{};
        md.n_groups = mgroups;
        md.elem_group = mgroup.data();
        md.direction = mdir.data();
        md.stiffness = mk.data();
        check_status(sors_set_muscle(ctx_, &md), ctx_);
        muscleGroups_ = mgroups;
    }
    if (!pgroups.empty()) {
        sors_pressure_desc pd{};
        pd.n_faces = (int)pgroups.size();
        pd.faces = pfaces.data();
        pd.groups = pgroups.data();
        pd.n_groups = pressureGroups_;
        check_status(sors_set_pressure(ctx_, &pd), ctx_);
    }
    if (hasContact_) {
        sors_contact_desc cd{};
        cd.n_planes = (int)(pp.size() / 3);
        cd.plane_point = pp.data();
        cd.plane_normal = pn.data();
        cd.n_disks = (int)dr.size();
        cd.disk_center = dc.data();
        cd.disk_normal = dn.data();
        cd.disk_radius = dr.data();
        cd.n_surface = (int)surf.size();
        cd.surface_nodes = surf.data();
        auto opt = [&](const char* key, double def) {
            return constraintParameterList.has(key) ? constraintParameterList.get_value(key)(0, 0) : def;
        };
        cd.stiffness = opt("contactStiffness", 1e5);
        cd.friction = opt("frictionCoefficient", 0.0);
        cd.eps_v = opt("frictionVelocity", 1e-3);
        check_status(sors_set_contact(ctx_, &cd), ctx_);
    }
    sors_options o;
    sors_default_options(&o);
    o.precision = settings.precision == "fp64" ? SORS_FP64 : SORS_MIXED;
    o.pcg_rtol = settings.pcgTolerance;
    o.predict = settings.predictor ? 2 : 0;
    o.coarse_nodes = settings.coarseNodes;
    o.coarse_every = settings.coarseEvery;
    check_status(sors_set_options(ctx_, &o), ctx_);
}

template <int E>
void Energy<E>::set_disk(const double center[3], const double normal[3], double radius, int diskIdx) {
    check_status(sors_set_disk(ctx_, diskIdx, center, normal, radius), ctx_);
}

template <int E>
Energy<E>::~Energy() {
    sors_destroy(ctx_);
}

template <int E>
void Energy<E>::set_initial_velocity(const Mat& v) {
    if (v.rows != numVertices_ || v.cols != 3) throw std::invalid_argument("velocities must be N x 3");
    auto r = v.row_major();
    check_status(sors_set_initial_velocity(ctx_, r.data()), ctx_);
}

template <int E>
void Energy<E>::set_initial_deformation(const Mat& x) {
    if (x.rows != numVertices_ || x.cols != 3) throw std::invalid_argument("vertices must be N x 3");
    auto r = x.row_major();
    check_status(sors_set_initial_deformation(ctx_, r.data()), ctx_);
}

template <int E>
std::vector<double> Energy<E>::get_vPrev() const {
    std::vector<double> v(3 * (size_t)numVertices_);
    check_status(sors_get_state(ctx_, nullptr, v.data()), ctx_);
    return v;
}

// VertexForce actuation: Params "vertexForce" (N x 3) -> 3N node-major
// (energy.cpp:282-293, vertexForce.cpp:4-8)
template <int E>
void Energy<E>::apply_actuation(const Params& act) {
    // "softconMuscle": one activation per muscle group (element.h:52,
    // softconMuscle.cpp:29 actuation.coeff(muscleGroup_))
    if (muscleGroups_ > 0 && act.has("softconMuscle")) {
        auto a = act.get_value("softconMuscle").row_major();
        if ((int)a.size() < muscleGroups_)
            throw std::invalid_argument("softconMuscle actuation needs one value per muscle group");
        check_status(sors_set_muscle_activation(ctx_, a.data()), ctx_);
    }
    // "pressure": one value per surface group (pressureForce.cpp:34, act.row(group))
    if (pressureGroups_ > 0 && act.has("pressure")) {
        auto p = act.get_value("pressure").row_major();  // transpose().reshaped()
        if ((int)p.size() < pressureGroups_)
            throw std::invalid_argument("pressure actuation needs one value per surface group");
        check_status(sors_set_pressure_activation(ctx_, p.data()), ctx_);
    }
    bool vf = false;
    for (const auto& f : forceTypesList_) vf |= (f == "vertexForce");
    if (!vf) return;
    if (!act.has("vertexForce")) {
        check_status(sors_set_external_force(ctx_, nullptr), ctx_);
        return;
    }
    auto f = act.get_value("vertexForce").row_major();
    if ((int)f.size() != 3 * numVertices_) throw std::invalid_argument("vertexForce actuation must be N x 3");
    check_status(sors_set_external_force(ctx_, f.data()), ctx_);
}

template <int E>
double Energy<E>::compute_energy(const std::vector<double>& q, double dt, const Params& actuation) {
    apply_actuation(actuation);
    double e = 0;
    check_status(sors_energy(ctx_, q.data(), dt, &e), ctx_);
    return e;
}

template <int E>
std::vector<double> Energy<E>::compute_gradient(const std::vector<double>& q, double dt, const Params& actuation) {
    apply_actuation(actuation);
    std::vector<double> g(q.size());
    check_status(sors_gradient(ctx_, q.data(), dt, g.data()), ctx_);
    return g;
}

template <int E>
SparseCSR Energy<E>::compute_hessian(const std::vector<double>& q, double dt, const Params& actuation) {
    if ((int)q.size() != 3 * numVertices_) throw std::invalid_argument("q must have 3N entries");
    apply_actuation(actuation);
    const long long nnz = sors_pattern_nnz(ctx_, dt);
    if (nnz < 0) check_status((int)nnz, ctx_);
    SparseCSR H;
    H.rows = H.cols = 3L * numVertices_;
    H.outer.resize((size_t)H.rows + 1);
    H.inner.resize((size_t)nnz);
    H.values.resize((size_t)nnz);
    check_status(sors_get_pattern(ctx_, dt, H.outer.data(), H.inner.data()), ctx_);
    check_status(sors_hessian(ctx_, q.data(), dt, SORS_FP64, 0, H.values.data()), ctx_);
    return H;
}

template <int E>
void Energy<E>::update_state(const std::vector<double>& q, double dt) {
    if ((int)q.size() != 3 * numVertices_) throw std::invalid_argument("q must have 3N entries");
    check_status(sors_update_state(ctx_, q.data(), dt), ctx_);
}

// ---------------------------------------------------------------- Solver
template <int E>
std::vector<double> Solver<E>::minimize(int method, const std::vector<double>& guess, Energy<E>& en, double dt,
                                        const Params& act) const {
    if (!(dt > 0.0)) throw std::invalid_argument("dt must be positive");
    if ((int)guess.size() != 3 * en.num_vertices()) throw std::invalid_argument("initialGuess must have 3N entries");
    en.apply_actuation(act);
    std::vector<double> q = guess;
    sors_stats st{};
    check_status(sors_minimize(en.ctx(), q.data(), dt, method, &st), en.ctx());
    last_stats = st;
    if (method == SORS_MINIMIZE_FROZENHESSIAN) {
        // solver.cpp:290-292: the frozen iteration reports, not raises, a miss
        sors_options o{};
        check_status(sors_get_options(en.ctx(), &o), en.ctx());
        if (!(st.dxnorm < o.dx_tol) && !(o.g_tol > 0 && st.gnorm < o.g_tol))
            std::cout << "Frozen Newton's method did not converge" << std::endl;
    }
    return q;
}

template <int E>
std::vector<double> Solver<E>::minimize_newton(const std::vector<double>& guess, Energy<E>& en, double dt,
                                               const Params& act) const {
    return minimize(SORS_MINIMIZE_NEWTON, guess, en, dt, act);
}

template <int E>
std::vector<double> Solver<E>::minimize_newton_linesearch(const std::vector<double>& guess, Energy<E>& en, double dt,
                                                          const Params& act) const {
    return minimize(SORS_MINIMIZE_NEWTON_LINESEARCH, guess, en, dt, act);
}

template <int E>
std::vector<double> Solver<E>::minimize_frozenhessian(const std::vector<double>& guess, Energy<E>& en, double dt,
                                                      const Params& act) const {
    return minimize(SORS_MINIMIZE_FROZENHESSIAN, guess, en, dt, act);
}

template <int E>
std::vector<double> Solver<E>::minimize_SQP(const std::vector<double>& guess, Energy<E>& en, double dt,
                                            const Params& act) const {
    return minimize(en.sqp_method(), guess, en, dt, act);
}

template <int E>
double Solver<E>::calculate_CFL_timestep(const std::vector<double>& solution, Energy<E>& en, double dt,
                                         double cflConstant) const {
    if ((int)solution.size() != 3 * en.num_vertices()) throw std::invalid_argument("solution must have 3N entries");
    double out = dt;
    check_status(sors_cfl_timestep(en.ctx(), solution.data(), dt, cflConstant, &out), en.ctx());
    return out;
}

// Solver::step (solver.cpp:5-44)
template <int E>
std::vector<double> Solver<E>::step(const std::vector<double>& solution, Energy<E>& en, const Params& act, double dt,
                                    int substeps) const {
    if (dt == 0.0) dt = en.settings_.dt;
    if (substeps == 0) substeps = en.settings_.substeps;
    if (!(dt > 0.0)) throw std::invalid_argument("dt must be positive (settings.dt is not set)");
    const bool frozen = en.settings_.solverMethod == "minimize_frozenhessian";
    if (en.settings_.solverMethod != "minimize_newton" && en.settings_.solverMethod != "minimize_newton_linesearch" &&
        !frozen && en.settings_.solverMethod != "minimize_SQP")
        throw std::runtime_error("Solver method not recognized");
    if ((int)solution.size() != 3 * en.num_vertices()) throw std::invalid_argument("solution must have 3N entries");
    if (en.settings_.CFLtimeSteppingFlag) {
        // solver.cpp:16-20: Courant number 1, from the positions the step starts at
        double dtc = dt;
        check_status(sors_cfl_timestep(en.ctx(), solution.data(), dt, 1.0, &dtc), en.ctx());
        substeps = (int)std::ceil(dt / dtc);
    }
    en.apply_actuation(act);
    std::vector<double> q = solution;
    sors_stats st{};
    check_status(sors_step(en.ctx(), q.data(), dt, substeps, &st), en.ctx());
    last_stats = st;
    if (frozen) {
        // solver.cpp:290-292: the frozen iteration reports, not raises, a miss
        sors_options o{};
        check_status(sors_get_options(en.ctx(), &o), en.ctx());
        if (!(st.dxnorm < o.dx_tol) && !(o.g_tol > 0 && st.gnorm < o.g_tol))
            std::cout << "Frozen Newton's method did not converge" << std::endl;
    }
    return q;
}

template <int E>
std::vector<double> Solver<E>::step(const std::vector<double>& solution, Energy<E>& en, double dt,
                                    int substeps) const {
    Params act;
    return step(solution, en, act, dt, substeps);
}

template class Energy<4>;
template class Energy<8>;
template class Solver<4>;
template class Solver<8>;

}  // namespace sors_host
