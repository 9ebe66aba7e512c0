// sors_elements.cpp -- element plugins and element classes over one-element
// device contexts (see sors_elements.h).
#include "sors_elements.h"

#include <cmath>
#include <numeric>
#include <stdexcept>

#include "sors_io.h"

namespace sors_host {

// ------------------------------------------------------------ DeviceElement
DeviceElement::DeviceElement(int E, const Mat& X, double density, const DeviceElementDesc& d)
    : E_(E), muscle_(d.muscleGroup >= 0) {
    if (X.rows != E || X.cols != 3) throw std::invalid_argument("element vertices must be E x 3");
    std::vector<double> rest = X.row_major();
    std::vector<int> conn(E);
    std::iota(conn.begin(), conn.end(), 0);
    // a muscle evaluated alone sits on a zero-stiffness Neo-Hookean element
    const int material = d.material >= 0 ? d.material : SORS_NEOHOOKEAN;
    double Y = d.youngs, nu = d.poisson, rho = density;
    sors_system_desc s{};
    s.n_nodes = E;
    s.n_elems = 1;
    s.nodes_per_elem = E;
    s.rest = rest.data();
    s.conn = conn.data();
    s.youngs = &Y;
    s.poisson = &nu;
    s.density = &rho;
    s.material = material;
    s.scheme = SORS_BACKWARD_EULER;
    s.method = SORS_MINIMIZE_NEWTON;
    check_status(sors_create(&s, &ctx_), nullptr);
    if (muscle_) {
        int group = 0;
        double k = d.muscleStiffness;
        std::array<double, 3> dir = d.muscleDirection;
        sors_muscle_desc m{};
        m.n_groups = 1;
        m.elem_group = &group;
        m.direction = dir.data();
        m.stiffness = &k;
        const int rc = sors_set_muscle(ctx_, &m);
        if (rc != SORS_OK) {
            std::string msg = sors_last_error(ctx_);
            sors_destroy(ctx_);
            ctx_ = nullptr;
            throw std::runtime_error(msg);
        }
    }
}

DeviceElement::~DeviceElement() {
    if (ctx_) sors_destroy(ctx_);
}

void DeviceElement::eval(const Mat& vertices, double activation, double* energy, std::vector<double>* grad,
                         Mat* hess) {
    if (vertices.rows != E_ || vertices.cols != 3) throw std::invalid_argument("element vertices must be E x 3");
    if (muscle_) check_status(sors_set_muscle_activation(ctx_, &activation), ctx_);
    const std::vector<double> q = vertices.row_major();
    const int nd = 3 * E_;
    double e = 0.0;
    std::vector<double> g(grad ? nd : 0), h(hess ? (size_t)nd * nd : 0);
    check_status(sors_element_eval(ctx_, q.data(), SORS_FP64, 0, &e, grad ? g.data() : nullptr,
                                   hess ? h.data() : nullptr),
                 ctx_);
    if (energy) *energy = e;
    if (grad) *grad = std::move(g);
    if (hess) {
        *hess = Mat(nd, nd);
        for (int r = 0; r < nd; ++r)
            for (int c = 0; c < nd; ++c) (*hess)(r, c) = h[(size_t)r * nd + c];
    }
}

// ------------------------------------------------------------ ElementEnergy
template <int E>
DeviceElementDesc ElementEnergy<E>::alone() const {
    DeviceElementDesc d;
    describe(d);
    return d;
}

template <int E>
DeviceElement& ElementEnergy<E>::device() const {
    if (!dev_) dev_ = std::make_unique<DeviceElement>(E, undeformed_, density_, alone());
    return *dev_;
}

template <int E>
double ElementEnergy<E>::compute_energy(const Mat& vertices, const std::vector<Mat>&, const Mat& actuation) const {
    double e = 0.0;
    device().eval(vertices, activation(actuation), &e, nullptr, nullptr);
    return e / volume_;
}

template <int E>
std::vector<double> ElementEnergy<E>::compute_gradient(const Mat& vertices, const std::vector<Mat>&,
                                                       const Mat& actuation) const {
    std::vector<double> g;
    device().eval(vertices, activation(actuation), nullptr, &g, nullptr);
    for (double& v : g) v /= volume_;
    return g;
}

template <int E>
Mat ElementEnergy<E>::compute_hessian(const Mat& vertices, const std::vector<Mat>&, const Mat& actuation) const {
    Mat h;
    device().eval(vertices, activation(actuation), nullptr, nullptr, &h);
    for (double& v : h.data) v /= volume_;
    return h;
}

template <int E>
void NeoHookeanEnergy<E>::describe(DeviceElementDesc& d) const {
    d.material = SORS_NEOHOOKEAN;
    d.youngs = youngs_;
    d.poisson = poisson_;
}

template <int E>
void StableNeoHookeanEnergy<E>::describe(DeviceElementDesc& d) const {
    d.material = SORS_STABLE_NEOHOOKEAN;
    d.youngs = this->youngs_;
    d.poisson = this->poisson_;
}

template <int E>
void SoftconMuscleEnergy<E>::describe(DeviceElementDesc& d) const {
    d.muscleGroup = group_;
    d.muscleStiffness = k_;
    d.muscleDirection = dir_;
}

// softconMuscle.cpp:29: actuation.coeff(muscleGroup_) of the "softconMuscle" entry
template <int E>
double SoftconMuscleEnergy<E>::activation(const Mat& actuation) const {
    if (actuation.size() == 0) return 0.0;
    const std::vector<double> a = actuation.row_major();
    if (group_ >= (int)a.size()) throw std::invalid_argument("softconMuscle actuation needs one value per muscle group");
    return a[(size_t)group_];
}

// ------------------------------------------------------------ registry
template <int E>
ElementEnergyRegistry<E>::ElementEnergyRegistry() {
    auto scalar = [](const Params& p, const char* key) { return p.get_value(key)(0, 0); };
    register_energy("neohookean", [scalar](const Params& p, const Mat& X, double rho, double V0) {
        return std::unique_ptr<ElementEnergy<E>>(
            new NeoHookeanEnergy<E>(X, rho, V0, scalar(p, "youngsModulus"), scalar(p, "poissonsRatio")));
    });
    register_energy("stableneohookean", [scalar](const Params& p, const Mat& X, double rho, double V0) {
        return std::unique_ptr<ElementEnergy<E>>(
            new StableNeoHookeanEnergy<E>(X, rho, V0, scalar(p, "youngsModulus"), scalar(p, "poissonsRatio")));
    });
    register_energy("softconMuscle", [scalar](const Params& p, const Mat& X, double rho, double V0) {
        const int group = (int)scalar(p, "muscleGroup");
        if (group < 0) throw std::invalid_argument("muscleGroup must be >= 0");
        const std::vector<double> dir = p.get_value("softconDirection").row_major();  // reshaped<RowMajor>()
        if (dir.size() != 3) throw std::invalid_argument("softconDirection must have 3 entries");
        return std::unique_ptr<ElementEnergy<E>>(new SoftconMuscleEnergy<E>(
            X, rho, V0, group, scalar(p, "softconStiffness"), {{dir[0], dir[1], dir[2]}}));
    });
}

template <int E>
ElementEnergyRegistry<E>& ElementEnergyRegistry<E>::instance() {
    static ElementEnergyRegistry<E> r;
    return r;
}

template <int E>
std::unique_ptr<ElementEnergy<E>> ElementEnergyRegistry<E>::create(const std::string& name, const Params& p,
                                                                   const Mat& X, double density, double V0) const {
    auto it = factories_.find(name);
    if (it == factories_.end()) {
        std::string known;
        for (const auto& kv : factories_) known += (known.empty() ? "\"" : ", \"") + kv.first + "\"";
        throw std::runtime_error("Unknown energy type: " + name + " (registered: " + known + ")");
    }
    return it->second(p, X, density, V0);
}

template <int E>
std::vector<std::string> ElementEnergyRegistry<E>::names() const {
    std::vector<std::string> out;
    for (const auto& kv : factories_) out.push_back(kv.first);
    return out;
}

// ------------------------------------------------------------ Element
template <int E>
Element<E>::Element(const Params& p, const Mat& X, const std::set<std::string>& energies, double V0, int nq)
    : mass_(p.get_value("density")(0, 0) * V0),
      undeformedVertices_(X),
      undeformedDensity_(p.get_value("density")(0, 0)),
      undeformedVolume_(V0),
      nquadpoints_(nq) {
    if (X.rows != E || X.cols != 3) throw std::invalid_argument("element vertices must be E x 3");
    // std::set order, as the reference iterates it
    for (const auto& name : energies) {
        elementEnergies_.push_back(
            ElementEnergyRegistry<E>::instance().create(name, p, X, undeformedDensity_, undeformedVolume_));
        elementEnergiesString_.push_back(name);
    }
}

template <int E>
Element<E>::~Element() = default;

template <int E>
DeviceElementDesc Element<E>::device_description() const {
    DeviceElementDesc d;
    for (const auto& en : elementEnergies_) {
        DeviceElementDesc one;
        en->describe(one);
        if (one.material >= 0) {
            if (d.material >= 0)
                throw std::runtime_error("the device path takes one elastic energy per element");
            d.material = one.material;
            d.youngs = one.youngs;
            d.poisson = one.poisson;
        }
        if (one.muscleGroup >= 0) {
            d.muscleGroup = one.muscleGroup;
            d.muscleStiffness = one.muscleStiffness;
            d.muscleDirection = one.muscleDirection;
        }
    }
    return d;
}

template <int E>
double Element<E>::element_activation(const Params& actuation) const {
    for (const auto& en : elementEnergies_)
        if (auto* m = dynamic_cast<const SoftconMuscleEnergy<E>*>(en.get())) {
            if (!actuation.has("softconMuscle")) return 0.0;
            const std::vector<double> a = actuation.get_value("softconMuscle").row_major();
            if (m->muscle_group() >= (int)a.size())
                throw std::invalid_argument("softconMuscle actuation needs one value per muscle group");
            return a[(size_t)m->muscle_group()];
        }
    return 0.0;
}

template <int E>
double Element<E>::compute_energy(const Mat& vertices, const Params& actuation) const {
    if (!dev_) dev_ = std::make_unique<DeviceElement>(E, undeformedVertices_, undeformedDensity_, device_description());
    double e = 0.0;
    dev_->eval(vertices, element_activation(actuation), &e, nullptr, nullptr);
    return e;
}

template <int E>
std::vector<double> Element<E>::compute_gradient(const Mat& vertices, const Params& actuation) const {
    if (!dev_) dev_ = std::make_unique<DeviceElement>(E, undeformedVertices_, undeformedDensity_, device_description());
    std::vector<double> g;
    dev_->eval(vertices, element_activation(actuation), nullptr, &g, nullptr);
    return g;
}

template <int E>
Mat Element<E>::compute_hessian(const Mat& vertices, const Params& actuation) const {
    if (!dev_) dev_ = std::make_unique<DeviceElement>(E, undeformedVertices_, undeformedDensity_, device_description());
    Mat h;
    dev_->eval(vertices, element_activation(actuation), nullptr, nullptr, &h);
    return h;
}

template <int E>
double Element<E>::compute_specific_energy(const Mat& vertices, const Params& actuation,
                                           const std::string& energyType) const {
    for (size_t i = 0; i < elementEnergiesString_.size(); ++i) {
        if (elementEnergiesString_[i] != energyType) continue;
        const ElementEnergy<E>& en = *elementEnergies_[i];
        Mat act;
        if (en.actuationFlag_ && actuation.has(en.get_energy_name())) act = actuation.get_value(en.get_energy_name());
        return en.compute_energy(vertices, {}, act) * undeformedVolume_;
    }
    return 0.0;
}

template <int E>
void Element<E>::update_state(const Mat& vertices, double dt) {
    for (auto& en : elementEnergies_) en->update_state(vertices, dt, {});
}

template <int E>
double Element<E>::compute_minimal_edge_length(const Mat& v) const {
    double best = -1.0;
    for (int i = 0; i < E; ++i)
        for (int j = i + 1; j < E; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += (v(i, k) - v(j, k)) * (v(i, k) - v(j, k));
            const double len = std::sqrt(s);
            if (best < 0.0 || len < best) best = len;
        }
    return best;
}

// ------------------------------------------------------------ Tetrahedron / Hexahedron
double Tetrahedron::volume_of(const Mat& v) {
    double ab[3], ac[3], ad[3];
    for (int k = 0; k < 3; ++k) {
        ab[k] = v(1, k) - v(0, k);
        ac[k] = v(2, k) - v(0, k);
        ad[k] = v(3, k) - v(0, k);
    }
    const double cx = ac[1] * ad[2] - ac[2] * ad[1], cy = ac[2] * ad[0] - ac[0] * ad[2],
                 cz = ac[0] * ad[1] - ac[1] * ad[0];
    return std::abs(ab[0] * cx + ab[1] * cy + ab[2] * cz) / 6.0;
}

Tetrahedron::Tetrahedron(const Params& p, const Mat& X, const std::set<std::string>& energies)
    : Element<4>(p, X, energies, volume_of(X), 1) {}

double Hexahedron::volume_of(const Mat& v) {
    if (v.rows != 8 || v.cols != 3) throw std::invalid_argument("hexahedron vertices must be 8 x 3");
    std::vector<int> local(8);
    std::iota(local.begin(), local.end(), 0);
    const std::vector<int> tets = IO::hex_to_5_tets(local);
    double vol = 0.0;
    for (size_t t = 0; t < tets.size() / 4; ++t) {
        Mat tv(4, 3);
        for (int r = 0; r < 4; ++r)
            for (int k = 0; k < 3; ++k) tv(r, k) = v(tets[4 * t + r], k);
        vol += Tetrahedron::volume_of(tv);
    }
    return vol;
}

Hexahedron::Hexahedron(const Params& p, const Mat& X, const std::set<std::string>& energies)
    : Element<8>(p, X, energies, volume_of(X), 8) {}

template class ElementEnergy<4>;
template class ElementEnergy<8>;
template class NeoHookeanEnergy<4>;
template class NeoHookeanEnergy<8>;
template class StableNeoHookeanEnergy<4>;
template class StableNeoHookeanEnergy<8>;
template class SoftconMuscleEnergy<4>;
template class SoftconMuscleEnergy<8>;
template class ElementEnergyRegistry<4>;
template class ElementEnergyRegistry<8>;
template class Element<4>;
template class Element<8>;

}  // namespace sors_host
