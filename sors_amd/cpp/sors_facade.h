// sors_facade.h -- the reference's host interfaces (Params, SimulationSettings,
// Energy<3,E>, Solver<3,E>) over the device C ABI in include/sors_amd.h.
//
// Same names, argument meaning and error messages as
//   cpp/headers/params.h:30-175, cpp/headers/simulationSettings.h:18-38,
//   cpp/headers/energy.h:73-205, cpp/headers/solver.h:27-95,
// with a small dense column-major matrix type standing in for Eigen::MatrixXd
// (Eigen is not a dependency).  Energies, constraints and forces outside the
// device path raise std::runtime_error instead of the reference's exit(1).
#pragma once

#include <cmath>
#include <iostream>
#include <limits>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "sors_amd.h"

namespace sors_host {

// column-major dense matrix (the role of Eigen::MatrixXd in the reference)
struct Mat {
    long rows = 0, cols = 0;
    std::vector<double> data;
    Mat() = default;
    Mat(long r, long c, double v = 0.0) : rows(r), cols(c), data((size_t)(r * c), v) {}
    double& operator()(long r, long c) { return data[(size_t)(c * rows + r)]; }
    double operator()(long r, long c) const { return data[(size_t)(c * rows + r)]; }
    long size() const { return rows * cols; }
    // row-major flattening, the reference's .reshaped<RowMajor>() / .transpose().reshaped()
    std::vector<double> row_major() const {
        std::vector<double> out((size_t)size());
        for (long r = 0; r < rows; ++r)
            for (long c = 0; c < cols; ++c) out[(size_t)(r * cols + c)] = (*this)(r, c);
        return out;
    }
};

// compressed sparse rows (the role of Eigen::SparseMatrix<double> returned by
// Energy::compute_hessian); coeff() is 0 for an entry not stored
struct SparseCSR {
    long rows = 0, cols = 0;
    std::vector<int> outer;     // rows + 1
    std::vector<int> inner;     // column of each stored value
    std::vector<double> values;
    long nonZeros() const { return (long)values.size(); }
    double coeff(long r, long c) const {
        for (int k = outer[(size_t)r]; k < outer[(size_t)r + 1]; ++k)
            if (inner[(size_t)k] == c) return values[(size_t)k];
        return 0.0;
    }
};

template <int E>
class Element;

// Key-value parameter store (cpp/headers/params.h:30-175)
class Params {
public:
    Params() = default;
    Params(const std::vector<std::string>& keys, const std::vector<Mat>& values);
    Params(const std::vector<std::string>& keys, const std::vector<double>& values);
    Mat get_value(const std::string& key) const;
    bool has(const std::string& key) const { return dict_.count(key) != 0; }
    void set_param(const std::string& key, const Mat& value);
    void set_param(const std::string& key, const std::vector<double>& values);
    void set_param(const std::string& key, double value);
    void set_param(const std::string& filename);  // CSV: key,value[,value...]
    const std::map<std::string, Mat>& dict() const { return dict_; }

private:
    std::map<std::string, Mat> dict_;
};

// Run settings (cpp/headers/simulationSettings.h:18-38), same defaults
struct SimulationSettings {
    std::string description;
    int verbose = 1;
    std::string meshFilePath;
    std::string outputFolder;
    std::string meshFileName;
    std::string meshType;
    double dt = std::numeric_limits<double>::quiet_NaN();
    int numSteps = -1;
    int substeps = 1;
    bool CFLtimeSteppingFlag = false;
    std::string timeSteppingScheme = "crank_nicolson";
    std::string solverMethod = "minimize_newton";
    int numThreads = 1;
    // device-path additions (not in the reference)
    std::string precision = "mixed";  // "mixed" | "fp64"
    double pcgTolerance = 1e-6;
    bool predictor = false;           // start Newton from the extrapolated state instead of the given solution
                                      // (the same minimiser).  Off: Solver::step starts Newton at the caller's
                                      // solution as the reference does (solver.cpp:22-40); the C ABI turns
                                      // it on by default (sors_default_options: predict = 2)
    int coarseNodes = 2048;           // > 0: two-level PCG (rigid-body coarse correction on aggregates of this
                                      // size); 0: block-Jacobi alone
    int coarseEvery = 16;             // coarse operator refreshed on every k-th tangent (as sors_default_options)
    std::string contactModel = "constraints";  // "constraints": minimize_SQP with planeContact / diskContact as
                                               // the reference's hard constraints; "penalty": this build's penalty
                                               // + Coulomb friction under Newton (contactStiffness,
                                               // frictionCoefficient, frictionVelocity)
    void verbose_print() const;
};

// Energy<3,E> over a device context (cpp/headers/energy.h:73-205)
template <int E>
class Energy {
public:
    Energy(const Mat& undeformedVertices, const std::vector<int>& eleIdxRowMajor, const SimulationSettings& settings,
           const double gravAcceleration[3], double dampingAlpha,
           const std::vector<std::set<std::string>>& elementEnergiesList,
           const std::vector<Params>& elementParameterList, const std::vector<std::string>& constraintTypesList,
           const Params& constraintParameterList, const std::vector<std::string>& forceTypesList,
           const Params& forceParameterList);
    ~Energy();
    Energy(const Energy&) = delete;
    Energy& operator=(const Energy&) = delete;

    void set_initial_velocity(const Mat& velocities);      // N x 3
    void set_initial_deformation(const Mat& vertices);     // N x 3
    std::vector<double> get_vPrev() const;
    double compute_energy(const std::vector<double>& q, double dt, const Params& actuation);
    std::vector<double> compute_gradient(const std::vector<double>& q, double dt, const Params& actuation);
    // energy.cpp:312-391: element Hessians (raw, not SPD-projected), force
    // terms, the integrator's mass / damping diagonal and the Neumann rows of
    // minimize_newton; FP64 values on the reference's CSR pattern
    SparseCSR compute_hessian(const std::vector<double>& q, double dt, const Params& actuation);
    // energy.cpp:539-557: the integrator's state (q_prev, v_prev, g_prev) from q;
    // the element plugins here keep no state of their own
    void update_state(const std::vector<double>& q, double dt);
    void apply_actuation(const Params& actuation);

    // DiskContact::set_disk (cpp/headers/constraints/diskContact.h:72-80)
    void set_disk(const double center[3], const double normal[3], double radius, int diskIdx);
    bool has_contact() const { return hasContact_; }

    sors_ctx* ctx() const { return ctx_; }
    int num_vertices() const { return numVertices_; }
    // the device method minimize_SQP resolves to: SORS_MINIMIZE_SQP, or with
    // contactModel "penalty" Newton on the penalty formulation
    int sqp_method() const { return sqpMethod_; }
    // energy.h: one Tetrahedron / Hexahedron per element, its energy plugins
    // built through the ElementEnergy registry; they evaluate on the device
    std::vector<std::unique_ptr<Element<E>>> elements_;
    SimulationSettings settings_;
    std::vector<std::string> forceTypesList_;
    std::vector<std::string> constraintTypesList_;

private:
    sors_ctx* ctx_ = nullptr;
    int numVertices_ = 0;
    bool hasContact_ = false;
    int muscleGroups_ = 0;
    int pressureGroups_ = 0;
    int sqpMethod_ = SORS_MINIMIZE_SQP;
};

// Non-owning handle on the diskContact entry of an Energy's constraint list
// (py_sors.cpp:231-261 get_disk_contact_constraint / DiskContact.set_disk)
template <int E>
struct DiskContactHandle {
    Energy<E>* energy;
    void set_disk(const double center[3], const double normal[3], double radius, int diskIdx = 0) {
        energy->set_disk(center, normal, radius, diskIdx);
    }
};

// Solver<3,E> (cpp/headers/solver.h:27-95)
template <int E>
class Solver {
public:
    Solver() = default;
    explicit Solver(const SimulationSettings& settings) : settings_(settings) {}
    SimulationSettings settings_;
    std::vector<double> step(const std::vector<double>& solution, Energy<E>& systemEnergy, const Params& act,
                             double dt = 0.0, int substeps = 0) const;
    std::vector<double> step(const std::vector<double>& solution, Energy<E>& systemEnergy, double dt = 0.0,
                             int substeps = 0) const;
    // solver.cpp:56-107, 201-266, 268-298: one minimisation of the step energy
    // from initialGuess; the energy's state is not updated (step = minimize +
    // update_state)
    std::vector<double> minimize_newton(const std::vector<double>& initialGuess, Energy<E>& systemEnergy, double dt,
                                        const Params& act) const;
    std::vector<double> minimize_newton_linesearch(const std::vector<double>& initialGuess, Energy<E>& systemEnergy,
                                                   double dt, const Params& act) const;
    std::vector<double> minimize_frozenhessian(const std::vector<double>& initialGuess, Energy<E>& systemEnergy,
                                               double dt, const Params& act) const;
    // solver.cpp:109-197: planeContact / diskContact / neumannBC as
    // constraints, one exactly solved QP per SQP iteration (contactModel
    // "penalty": Newton with line search on the penalty formulation, plain
    // Newton with a muscle, whose energy is not the potential of its gradient)
    std::vector<double> minimize_SQP(const std::vector<double>& initialGuess, Energy<E>& systemEnergy, double dt,
                                     const Params& act) const;
    // solver.cpp:300-331
    double calculate_CFL_timestep(const std::vector<double>& solution, Energy<E>& systemEnergy, double dt,
                                  double cflConstant) const;
    mutable sors_stats last_stats{};

private:
    std::vector<double> minimize(int method, const std::vector<double>& initialGuess, Energy<E>& systemEnergy,
                                 double dt, const Params& act) const;
};

// throws std::runtime_error carrying the reference's message for a status code
void check_status(int code, const sors_ctx* ctx);

}  // namespace sors_host
