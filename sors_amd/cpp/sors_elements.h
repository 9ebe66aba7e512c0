// sors_elements.h -- the reference's per-element plugin interface and element
// classes, host side (cpp/headers/energies/elementEnergy.h:16-114,
// cpp/headers/geometries/element.h, tetrahedron.h, hexahedron.h), over the
// device C ABI in include/sors_amd.h.
//
// An element energy plugin here is the host description of a device kernel
// family: its name (the string of Energy's elementEnergiesList), its
// per-element parameters, and what it sets in the device context (the
// elastic material, or a muscle entry).  The Energy constructor builds every
// element's plugins through a name registry, and the device context from
// their descriptions.  Per-element evaluation (Element::compute_energy /
// gradient / hessian, ElementEnergy::compute_*) runs on the device through a
// one-element context created on first use; no element math runs on the host.
#pragma once

#include <array>
#include <functional>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "sors_facade.h"

namespace sors_host {

// What one element's plugins set in the device context.
struct DeviceElementDesc {
    int material = -1;              // SORS_NEOHOOKEAN / SORS_STABLE_NEOHOOKEAN (elastic plugins)
    double youngs = 0.0, poisson = 0.0;
    int muscleGroup = -1;           // softconMuscle: group, stiffness, direction
    double muscleStiffness = 0.0;
    std::array<double, 3> muscleDirection{{0.0, 0.0, 0.0}};
};

// A one-element device context: evaluates the element's energy (x V0),
// gradient [3E] and Hessian [3E x 3E] at given vertex positions, in FP64 and
// without SPD projection (Element::compute_hessian is the raw Hessian,
// neoHookean.cpp:74-126).
class DeviceElement {
public:
    DeviceElement(int E, const Mat& undeformedVertices, double density, const DeviceElementDesc& d);
    ~DeviceElement();
    DeviceElement(const DeviceElement&) = delete;
    DeviceElement& operator=(const DeviceElement&) = delete;
    // vertices: E x 3; activation of the muscle (if any)
    void eval(const Mat& vertices, double activation, double* energy, std::vector<double>* grad, Mat* hess);

private:
    int E_;
    bool muscle_;
    sors_ctx* ctx_ = nullptr;
};

// ElementEnergy<3,E> (elementEnergy.h:16-114).  compute_* return the values
// per unit undeformed volume, as the reference's plugins do (the element
// multiplies by its volume); the deformation gradients of the reference's
// signature are formed on the device from the vertices, so the vector passed
// is not read and may be empty.
template <int E>
class ElementEnergy {
public:
    ElementEnergy(const Mat& undeformedVertices, double density, double undeformedVolume)
        : undeformed_(undeformedVertices), density_(density), volume_(undeformedVolume) {}
    virtual ~ElementEnergy() = default;
    virtual std::string get_energy_name() const = 0;
    // this plugin's share of the element's device description
    virtual void describe(DeviceElementDesc& d) const = 0;

    double compute_energy(const Mat& vertices, const std::vector<Mat>& deformationGradients, const Mat& actuation) const;
    std::vector<double> compute_gradient(const Mat& vertices, const std::vector<Mat>& deformationGradients,
                                         const Mat& actuation) const;
    Mat compute_hessian(const Mat& vertices, const std::vector<Mat>& deformationGradients, const Mat& actuation) const;
    // elementEnergy.h:93-99: stateless plugins keep the default no-op
    virtual void update_state(const Mat& vertices, double dt, const std::vector<Mat>& deformationGradients) {
        (void)vertices;
        (void)dt;
        (void)deformationGradients;
    }
    bool actuationFlag_ = false;

protected:
    // the activation this plugin reads from its actuation entry
    virtual double activation(const Mat& actuation) const {
        (void)actuation;
        return 0.0;
    }
    // the device description of this plugin alone (an elastic plugin alone,
    // or a muscle on a zero-stiffness Neo-Hookean element)
    DeviceElementDesc alone() const;

    Mat undeformed_;
    double density_, volume_;
    mutable std::unique_ptr<DeviceElement> dev_;
    DeviceElement& device() const;
};

// "neohookean" / "stableneohookean" (neoHookean.cpp, stableNeoHookean.cpp;
// Lame constants from Young's modulus and Poisson's ratio on the device)
template <int E>
class NeoHookeanEnergy : public ElementEnergy<E> {
public:
    NeoHookeanEnergy(const Mat& X, double density, double V0, double youngsModulus, double poissonsRatio)
        : ElementEnergy<E>(X, density, V0), youngs_(youngsModulus), poisson_(poissonsRatio) {}
    std::string get_energy_name() const override { return "neohookean"; }
    void describe(DeviceElementDesc& d) const override;

protected:
    double youngs_, poisson_;
};

template <int E>
class StableNeoHookeanEnergy : public NeoHookeanEnergy<E> {
public:
    using NeoHookeanEnergy<E>::NeoHookeanEnergy;
    std::string get_energy_name() const override { return "stableneohookean"; }
    void describe(DeviceElementDesc& d) const override;
};

// "softconMuscle" (softconMuscle.cpp:12-134): actuated by entry
// actuation("softconMuscle")(0, muscleGroup)
template <int E>
class SoftconMuscleEnergy : public ElementEnergy<E> {
public:
    SoftconMuscleEnergy(const Mat& X, double density, double V0, int muscleGroup, double stiffness,
                        const std::array<double, 3>& direction)
        : ElementEnergy<E>(X, density, V0), group_(muscleGroup), k_(stiffness), dir_(direction) {
        this->actuationFlag_ = true;
    }
    std::string get_energy_name() const override { return "softconMuscle"; }
    void describe(DeviceElementDesc& d) const override;
    int muscle_group() const { return group_; }

protected:
    double activation(const Mat& actuation) const override;
    int group_;
    double k_;
    std::array<double, 3> dir_;
};

// Name -> plugin factory: the string-to-class dispatch of the reference's
// element constructors (tetrahedron.cpp:55-95, hexahedron.cpp:24-64).
// Factories take the element's Params, undeformed vertices, density and
// volume.  register_energy adds (or replaces) a name.
template <int E>
class ElementEnergyRegistry {
public:
    using Factory = std::function<std::unique_ptr<ElementEnergy<E>>(const Params&, const Mat&, double, double)>;
    static ElementEnergyRegistry& instance();
    void register_energy(const std::string& name, Factory f) { factories_[name] = std::move(f); }
    bool has(const std::string& name) const { return factories_.count(name) != 0; }
    // throws std::runtime_error("Unknown energy type: ...") for a name not registered
    std::unique_ptr<ElementEnergy<E>> create(const std::string& name, const Params& p, const Mat& X, double density,
                                             double V0) const;
    std::vector<std::string> names() const;

private:
    ElementEnergyRegistry();
    std::map<std::string, Factory> factories_;
};

// Element<3,E> (element.h): undeformed geometry, mass and the element's
// energy plugins.  Energies and derivatives come from the device.
template <int E>
class Element {
public:
    Element(const Params& parameterMap, const Mat& undeformedVertices, const std::set<std::string>& energies,
            double undeformedVolume, int nquadpoints);
    virtual ~Element();
    Element(const Element&) = delete;
    Element& operator=(const Element&) = delete;

    // element.h:40-58 / tetrahedron.cpp:99-134 / hexahedron.cpp:68-103
    double compute_energy(const Mat& vertices, const Params& actuation) const;
    std::vector<double> compute_gradient(const Mat& vertices, const Params& actuation) const;
    Mat compute_hessian(const Mat& vertices, const Params& actuation) const;
    // element.h:61-76: one named energy of the element (0 when absent)
    double compute_specific_energy(const Mat& vertices, const Params& actuation, const std::string& energyType) const;
    // element.h:79-88
    void update_state(const Mat& vertices, double dt);
    // element.h:131-142: shortest distance between two vertices of the element
    double compute_minimal_edge_length(const Mat& vertices) const;
    double compute_mass() const { return mass_; }
    double get_volume(const Mat& vertices) const { return compute_volume(vertices); }
    double get_undeformed_volume() const { return undeformedVolume_; }
    Mat get_undeformed_vertices() const { return undeformedVertices_; }
    double get_density() const { return undeformedDensity_; }
    const std::vector<std::string>& energy_names() const { return elementEnergiesString_; }
    // the element's device description (all its plugins)
    DeviceElementDesc device_description() const;

    std::vector<std::unique_ptr<ElementEnergy<E>>> elementEnergies_;
    std::vector<std::string> elementEnergiesString_;

protected:
    virtual double compute_volume(const Mat& vertices) const = 0;
    double element_activation(const Params& actuation) const;
    double mass_;
    Mat undeformedVertices_;
    double undeformedDensity_;
    double undeformedVolume_;
    int nquadpoints_;
    mutable std::unique_ptr<DeviceElement> dev_;
};

// Tetrahedron (tetrahedron.h, one quadrature point) and Hexahedron
// (hexahedron.h, 2 x 2 x 2 Gauss points); vertices E x 3 in the reference's
// local order.
class Tetrahedron : public Element<4> {
public:
    Tetrahedron(const Params& parameterMap, const Mat& undeformedVertices, const std::set<std::string>& energies);
    // tetrahedron.cpp:137-149: |det| / 6
    static double volume_of(const Mat& vertices);

protected:
    double compute_volume(const Mat& vertices) const override { return volume_of(vertices); }
};

class Hexahedron : public Element<8> {
public:
    Hexahedron(const Params& parameterMap, const Mat& undeformedVertices, const std::set<std::string>& energies);
    // hexahedron.cpp:106-126: the five tetrahedra of the split, summed
    static double volume_of(const Mat& vertices);

protected:
    double compute_volume(const Mat& vertices) const override { return volume_of(vertices); }
};

template <int E>
using ElementType = typename std::conditional<E == 4, Tetrahedron, Hexahedron>::type;

}  // namespace sors_host
