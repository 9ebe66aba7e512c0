// C++ test of the host interfaces over the device (run by tests/test_gpu_cpp_facade.py).
//
// Builds two small beams through the reference's C++ surface -- Energy<3,E>,
// Solver<3,E>, the ElementEnergy plugins and the Element classes -- and
// prints one JSON object with every input it used and every value it got:
//   - Energy::compute_energy / compute_gradient / compute_hessian at a
//     perturbed state (the Python side checks them against the FP64 oracle);
//   - Element::compute_energy / gradient / hessian and the plugins' own
//     values for a few elements (against the oracle's element evaluation);
//   - Solver::minimize_* from a guess, and Solver::step against minimize +
//     Energy::update_state on a twin Energy (checked here: bit for bit);
//   - Energy::update_state's v_prev, calculate_CFL_timestep, and the
//     registry's error for an unknown energy name.
// Scenario "tet": tet4 Neo-Hookean, backward Euler, Newton + line search.
// Scenario "hex": hex8 Stable Neo-Hookean with SoftCon muscles on half the
// elements, Crank-Nicolson, plain Newton.
#include <cmath>
#include <cstdio>
#include <functional>
#include <memory>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "sors_elements.h"
#include "sors_facade.h"
#include "sors_io.h"

using namespace sors_host;

namespace {

std::string num(double v) {
    char b[40];
    std::snprintf(b, sizeof b, "%.17g", v);
    return b;
}

template <typename T>
std::string arr(const std::vector<T>& v) {
    std::ostringstream o;
    o << "[";
    for (size_t i = 0; i < v.size(); ++i) o << (i ? "," : "") << num((double)v[i]);
    o << "]";
    return o.str();
}

std::string mat(const Mat& m) {  // row-major nested list
    std::ostringstream o;
    o << "[";
    for (long r = 0; r < m.rows; ++r) {
        o << (r ? "," : "") << "[";
        for (long c = 0; c < m.cols; ++c) o << (c ? "," : "") << num(m(r, c));
        o << "]";
    }
    o << "]";
    return o.str();
}

struct Scenario {
    std::string name, meshType, scheme, method;
    int E = 4;
    Mat X;
    std::vector<int> conn;
    std::vector<std::set<std::string>> energies;
    std::vector<Params> params;
    std::vector<double> youngs, poisson, density;
    std::vector<int> muscleGroup;
    std::vector<double> muscleK, muscleDir, activation;
    std::vector<double> mask;
    double dt = 1e-3;
    double grav[3] = {0.0, 0.0, -9.81};
};

Scenario make(const std::string& name) {
    Scenario s;
    s.name = name;
    const bool tet = name == "tet";
    s.E = tet ? 4 : 8;
    s.meshType = tet ? "Tetrahedron" : "Hexahedron";
    s.scheme = tet ? "backward_euler" : "crank_nicolson";
    s.method = tet ? "minimize_newton_linesearch" : "minimize_newton";
    auto grid = IO::make_voxel_hex_grid(tet ? 6 : 5, 2, 2, 0.01, 0.01, 0.01);
    s.X = grid.first;
    s.conn = tet ? IO::hex_to_5_tets(grid.second) : grid.second;
    const int nel = (int)(s.conn.size() / s.E), N = (int)s.X.rows;
    s.mask.assign(3 * (size_t)N, 0.0);
    for (int v = 0; v < N; ++v)
        if (s.X(v, 0) < 1e-12)
            for (int k = 0; k < 3; ++k) s.mask[3 * (size_t)v + k] = 1.0;
    s.activation = {0.3, 0.7};
    for (int e = 0; e < nel; ++e) {
        const double Y = 1e5 * (1.0 + (e % 7) / 7.0), nu = 0.3 + 0.02 * (e % 5), rho = 1000.0 + 10.0 * (e % 3);
        s.youngs.push_back(Y);
        s.poisson.push_back(nu);
        s.density.push_back(rho);
        std::set<std::string> en;
        Params p(std::vector<std::string>{"density", "youngsModulus", "poissonsRatio"},
                 std::vector<double>{rho, Y, nu});
        int g = -1;
        double k = 0.0, d[3] = {0.0, 0.0, 0.0};
        if (tet) {
            en.insert("neohookean");
        } else {
            en.insert("stableneohookean");
            if (e % 2 == 0) {
                en.insert("softconMuscle");
                g = (e / 2) % 2;
                k = 2e4 * (1.0 + 0.5 * g);
                d[0] = 1.0;
                d[1] = 0.2 * g;
                d[2] = 0.1;
                p.set_param("muscleGroup", (double)g);
                p.set_param("softconStiffness", k);
                p.set_param("softconDirection", std::vector<double>{d[0], d[1], d[2]});
            }
        }
        s.muscleGroup.push_back(g);
        s.muscleK.push_back(k);
        for (int c = 0; c < 3; ++c) s.muscleDir.push_back(d[c]);
        s.energies.push_back(en);
        s.params.push_back(p);
    }
    return s;
}

std::unique_ptr<Energy<4>> energy4(const Scenario& s, const SimulationSettings& st) {
    Params cp;
    Mat m(1, (long)s.mask.size()), v(1, (long)s.mask.size());
    for (size_t i = 0; i < s.mask.size(); ++i) m(0, (long)i) = s.mask[i];
    cp.set_param("neumannBCmask", m);
    cp.set_param("neumannBCvalue", v);
    return std::make_unique<Energy<4>>(s.X, s.conn, st, s.grav, 0.0, s.energies, s.params,
                                       std::vector<std::string>{"neumannBC"}, cp, std::vector<std::string>{}, Params());
}

std::unique_ptr<Energy<8>> energy8(const Scenario& s, const SimulationSettings& st) {
    Params cp;
    Mat m(1, (long)s.mask.size()), v(1, (long)s.mask.size());
    for (size_t i = 0; i < s.mask.size(); ++i) m(0, (long)i) = s.mask[i];
    cp.set_param("neumannBCmask", m);
    cp.set_param("neumannBCvalue", v);
    return std::make_unique<Energy<8>>(s.X, s.conn, st, s.grav, 0.0, s.energies, s.params,
                                       std::vector<std::string>{"neumannBC"}, cp, std::vector<std::string>{}, Params());
}

template <int E>
std::string run(const Scenario& s, std::function<std::unique_ptr<Energy<E>>(const Scenario&, const SimulationSettings&)> mk) {
    SimulationSettings st;
    st.meshType = s.meshType;
    st.timeSteppingScheme = s.scheme;
    st.solverMethod = s.method;
    st.dt = s.dt;
    st.precision = "fp64";
    st.pcgTolerance = 1e-12;
    st.coarseNodes = 0;
    const int N = (int)s.X.rows, n = 3 * N;
    std::vector<double> X = s.X.row_major();
    // initial velocity: a bend about y, zero at the clamp
    Mat V0(N, 3);
    for (int v = 0; v < N; ++v) V0(v, 2) = 0.4 * s.X(v, 0) / 0.06 * (s.mask[3 * (size_t)v] == 0.0);
    Params act;
    if (E == 8) act.set_param("softconMuscle", s.activation);
    std::ostringstream o;
    o << "{\"meshType\":\"" << s.meshType << "\",\"scheme\":\"" << s.scheme << "\",\"method\":\"" << s.method
      << "\",\"dt\":" << num(s.dt) << ",\"X\":" << arr(X) << ",\"conn\":" << arr(s.conn)
      << ",\"youngs\":" << arr(s.youngs) << ",\"poisson\":" << arr(s.poisson) << ",\"density\":" << arr(s.density)
      << ",\"mask\":" << arr(s.mask) << ",\"muscleGroup\":" << arr(s.muscleGroup) << ",\"muscleK\":" << arr(s.muscleK)
      << ",\"muscleDir\":" << arr(s.muscleDir) << ",\"activation\":" << arr(s.activation)
      << ",\"v0\":" << arr(V0.row_major());

    auto en = mk(s, st);
    en->set_initial_velocity(V0);
    // a perturbed state (the clamp kept)
    std::vector<double> q(X);
    for (int i = 0; i < n; ++i)
        if (s.mask[(size_t)i] == 0.0) q[(size_t)i] += 4e-4 * std::sin(1.3 * i + 0.2);
    o << ",\"q\":" << arr(q);
    o << ",\"energy\":" << num(en->compute_energy(q, s.dt, act));
    o << ",\"gradient\":" << arr(en->compute_gradient(q, s.dt, act));
    const SparseCSR H = en->compute_hessian(q, s.dt, act);
    o << ",\"hessian\":{\"outer\":" << arr(H.outer) << ",\"inner\":" << arr(H.inner) << ",\"values\":" << arr(H.values)
      << "}";

    // elements: whole-element values and each plugin's own, at q
    o << ",\"elements\":[";
    const int nel = (int)en->elements_.size();
    const int pick[3] = {0, 1, nel - 1};
    for (int t = 0; t < 3; ++t) {
        const int e = pick[t];
        const Element<E>& el = *en->elements_[(size_t)e];
        Mat xe(E, 3);
        for (int a = 0; a < E; ++a)
            for (int k = 0; k < 3; ++k) xe(a, k) = q[3 * (size_t)s.conn[(size_t)E * e + a] + k];
        o << (t ? "," : "") << "{\"e\":" << e << ",\"energy\":" << num(el.compute_energy(xe, act))
          << ",\"gradient\":" << arr(el.compute_gradient(xe, act)) << ",\"hessian\":" << mat(el.compute_hessian(xe, act))
          << ",\"volume\":" << num(el.get_undeformed_volume()) << ",\"mass\":" << num(el.compute_mass())
          << ",\"min_edge\":" << num(el.compute_minimal_edge_length(xe)) << ",\"plugins\":{";
        for (size_t i = 0; i < el.elementEnergies_.size(); ++i) {
            const ElementEnergy<E>& pl = *el.elementEnergies_[i];
            Mat a;
            if (pl.actuationFlag_) a = act.get_value(pl.get_energy_name());
            o << (i ? "," : "") << "\"" << pl.get_energy_name() << "\":{\"energy\":" << num(pl.compute_energy(xe, {}, a))
              << ",\"specific\":" << num(el.compute_specific_energy(xe, act, pl.get_energy_name()))
              << ",\"gradient\":" << arr(pl.compute_gradient(xe, {}, a)) << "}";
        }
        o << "}}";
    }
    o << "]";

    // minimisers from the constant-velocity guess
    Solver<E> solver(st);
    std::vector<double> guess(X);
    const std::vector<double> v0 = V0.row_major();
    for (int i = 0; i < n; ++i) guess[(size_t)i] += s.dt * v0[(size_t)i];
    o << ",\"guess\":" << arr(guess);
    o << ",\"minimize_newton\":" << arr(solver.minimize_newton(guess, *en, s.dt, act));
    o << ",\"minimize_newton_linesearch\":" << arr(solver.minimize_newton_linesearch(guess, *en, s.dt, act));
    o << ",\"cfl\":" << num(solver.calculate_CFL_timestep(X, *en, s.dt, 1.0));

    // step == minimize + update_state, twice, on twin energies
    auto a = mk(s, st);
    auto b = mk(s, st);
    a->set_initial_velocity(V0);
    b->set_initial_velocity(V0);
    std::vector<double> qa = X, qb = X;
    bool same = true;
    for (int k = 0; k < 2; ++k) {
        qa = solver.step(qa, *a, act, s.dt);
        qb = s.method == "minimize_newton" ? solver.minimize_newton(qb, *b, s.dt, act)
                                           : solver.minimize_newton_linesearch(qb, *b, s.dt, act);
        b->update_state(qb, s.dt);
        same = same && qa == qb && a->get_vPrev() == b->get_vPrev();
    }
    o << ",\"step_equals_minimize_update\":" << (same ? "true" : "false");
    o << ",\"step2\":" << arr(qa) << ",\"vprev2\":" << arr(a->get_vPrev());

    // unknown plugin name
    std::string err;
    try {
        Scenario t = s;
        t.energies[0] = {"pseudostrainX"};
        mk(t, st);
    } catch (const std::exception& ex) {
        err = ex.what();
    }
    for (char& c : err)
        if (c == '"' || c == '\\') c = '\'';
    o << ",\"unknown_energy_error\":\"" << err << "\"";
    std::string names;
    for (const auto& nm : ElementEnergyRegistry<E>::instance().names()) names += (names.empty() ? "" : " ") + nm;
    o << ",\"registry\":\"" << names << "\"}";
    return o.str();
}

}  // namespace

int main() {
    try {
        const std::string t = run<4>(make("tet"), energy4);
        const std::string h = run<8>(make("hex"), energy8);
        std::printf("{\"tet\":%s,\"hex\":%s}\n", t.c_str(), h.c_str());
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "test_facade: %s\n", e.what());
        return 1;
    }
}
