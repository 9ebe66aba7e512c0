// sors_demos.h -- the demo entry point the reference binds in py_sors
// (cpp/py_sors.cpp:27, cpp/demos/demos.h:10), over the device path.
#pragma once

#include <string>

namespace sors_host {

// run_msd_timestepping (cpp/demos/03_msd_timestepping.cpp:16-212): the
// mass-spring-damper cylinder under gravity with its top clamped, Crank-
// Nicolson with CFL sub-stepping, a lateral push on the cube half between
// steps 105 and 125, one VTU per step, the PVD and msd_displacement.csv.
void run_msd_timestepping(const std::string& meshFilePath, const std::string& outputFolder, int numSteps);

}  // namespace sors_host
