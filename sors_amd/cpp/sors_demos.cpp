// sors_demos.cpp -- run_msd_timestepping over the façade (Energy<4>, Solver<4>,
// IO), following cpp/demos/03_msd_timestepping.cpp:16-212 step by step.
#include "sors_demos.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <fstream>
#include <iostream>
#include <stdexcept>

#include "sors_facade.h"
#include "sors_io.h"

namespace sors_host {

void run_msd_timestepping(const std::string& meshFilePath, const std::string& outputFolder, int numSteps) {
    // simulation settings (:19-29)
    SimulationSettings settings;
    settings.description =
        "Demo of a mass-spring damper with external forces under gravity and adaptive time stepping.";
    settings.meshFilePath = meshFilePath;
    settings.outputFolder = outputFolder;
    settings.meshFileName = meshFilePath.substr(meshFilePath.find_last_of('/') + 1);
    settings.dt = 1e-2;
    settings.numSteps = numSteps;
    settings.CFLtimeSteppingFlag = true;
    settings.timeSteppingScheme = "crank_nicolson";
    settings.solverMethod = "minimize_newton";
    settings.verbose_print();

    // mesh in mm, scaled to m (:33-35)
    Mat V;
    std::vector<int> T;
    IO::read_tet_MSH(meshFilePath, V, T);
    for (double& x : V.data) x *= 1e-3;
    const long N = V.rows, nel = (long)T.size() / 4;
    double lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = hi[k] = N ? V(0, k) : 0.0;
        for (long i = 1; i < N; ++i) {
            lo[k] = std::min(lo[k], V(i, k));
            hi[k] = std::max(hi[k], V(i, k));
        }
    }
    const double minX = lo[0], maxX = hi[0], minY = lo[1], maxY = hi[1], minZ = lo[2], maxZ = hi[2];

    // element centres (:47-57), Neo-Hookean everywhere (:59-69)
    std::vector<double> cz((size_t)nel, 0.0);
    std::vector<std::set<std::string>> energies((size_t)nel);
    std::vector<Params> params((size_t)nel);
    for (long e = 0; e < nel; ++e) {
        for (int j = 0; j < 4; ++j) cz[(size_t)e] += V(T[(size_t)(4 * e + j)], 2);
        cz[(size_t)e] /= 4.0;
        energies[(size_t)e].insert("neohookean");
        params[(size_t)e].set_param("density", 1070.);
        params[(size_t)e].set_param("youngsModulus", 12.5 * 263824.0);
        params[(size_t)e].set_param("poissonsRatio", 0.499);
    }

    // the top of the rod fixed (:71-90)
    std::vector<std::string> constraintTypes{"neumannBC"};
    Mat mask(3 * N, 1, 0.0), value(3 * N, 1, 0.0);
    for (long i = 0; i < N; ++i)
        if (V(i, 2) > maxZ - 1e-4)
            for (int k = 0; k < 3; ++k) mask(3 * i + k, 0) = 1.0;
    Params constraintParams;
    constraintParams.set_param("neumannBCmask", mask);
    constraintParams.set_param("neumannBCvalue", value);

    // time-dependent vertex forces (:92-99)
    std::vector<std::string> forceTypes{"vertexForce"};
    Params forceParams, actuation;
    Mat forces(N, 3, 0.0);
    actuation.set_param("vertexForce", forces);

    // cube (lower half) and rod vertices; element colours (:101-126)
    std::vector<int> cube((size_t)N, 0);
    for (long i = 0; i < N; ++i) cube[(size_t)i] = V(i, 2) <= (maxZ - minZ) / 2 ? 1 : 0;
    VisualOption colors;
    colors.name = "elementColor";
    colors.type = VisualType::ScalarElement;
    colors.data = Mat(1, nel, 0.0);
    for (long e = 0; e < nel; ++e) colors.data(0, e) = cz[(size_t)e] <= (maxZ - minZ) / 2 ? 0.0 : 1.0;

    // two opposite cube corners tracked for the centre of mass (:128-148)
    const double tol = 1e-3;
    long t1 = -1, t2 = -1;
    for (long i = 0; i < N; ++i) {
        if (std::abs(V(i, 0) - minX) < tol && std::abs(V(i, 1) - minY) < tol && std::abs(V(i, 2) - minZ) < tol) t1 = i;
        if (std::abs(V(i, 0) - maxX) < tol && std::abs(V(i, 1) - maxY) < tol && std::abs(V(i, 2) - maxZ / 2.0) < tol)
            t2 = i;
    }
    if (t1 < 0 || t2 < 0) throw std::runtime_error("run_msd_timestepping: tracking vertices not found in the mesh");
    std::vector<double> comHeight((size_t)std::max(numSteps, 1), 0.0), cubeForce((size_t)std::max(numSteps, 1), 0.0);
    comHeight[0] = (V(t1, 2) + V(t2, 2)) / 2.0;

    // system energy (mass-proportional damping 1) and solver (:150-153)
    const double grav[3] = {0.0, 0.0, -9.81};
    Energy<4> systemEnergy(V, T, settings, grav, 1.0, energies, params, constraintTypes, constraintParams, forceTypes,
                           forceParams);
    Solver<4> solver;

    const std::string base = settings.outputFolder + "/" + settings.meshFileName;
    IO::save_tet_VTU(base + "_0", V, T, {colors});
    std::vector<double> solution = V.row_major();
    auto begin = std::chrono::steady_clock::now();
    const double forceMagnitude = 0.5;
    for (int i = 1; i < settings.numSteps; ++i) {
        // lateral push on the cube between steps 105 and 125 (:165-176)
        if (i == 105) {
            for (long j = 0; j < N; ++j)
                if (cube[(size_t)j]) forces(j, 0) = forceMagnitude * 0.01;
            actuation.set_param("vertexForce", forces);
        }
        if (i == 125) actuation.set_param("vertexForce", Mat(N, 3, 0.0));
        solution = solver.step(solution, systemEnergy, actuation);
        comHeight[(size_t)i] = (solution[(size_t)(3 * t1 + 2)] + solution[(size_t)(3 * t2 + 2)]) / 2.0;
        const Mat f = actuation.get_value("vertexForce");
        cubeForce[(size_t)i] = std::sqrt(f(t1, 0) * f(t1, 0) + f(t1, 1) * f(t1, 1) + f(t1, 2) * f(t1, 2));
        Mat Vi(N, 3);
        for (long j = 0; j < N; ++j)
            for (int k = 0; k < 3; ++k) Vi(j, k) = solution[(size_t)(3 * j + k)];
        IO::save_tet_VTU(base + "_" + std::to_string(i), Vi, T, {colors});
        const auto ms =
            std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - begin).count();
        std::cout << "Time until simulation step " << i << ": " << ms << " ms" << std::endl;
    }
    const auto ms =
        std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - begin).count();
    if (settings.numSteps > 1)
        std::cout << "Average time per simulation step: " << ms / (settings.numSteps - 1) << " ms" << std::endl;
    IO::save_pvd(settings.meshFileName, settings.outputFolder, settings.numSteps, settings.dt);

    // step, time, centre-of-mass height, force on the tracked vertex (:200-209)
    std::ofstream file(outputFolder + "/msd_displacement.csv");
    if (!file.is_open()) {
        std::cerr << "Could not open file for writing" << std::endl;
        return;
    }
    file.precision(17);
    for (int i = 0; i < settings.numSteps; ++i)
        file << i << "," << i * settings.dt << "," << comHeight[(size_t)i] << "," << cubeForce[(size_t)i] << "\n";
    file.close();
    std::cout << "Simulation Sucessful" << std::endl;
}

}  // namespace sors_host
