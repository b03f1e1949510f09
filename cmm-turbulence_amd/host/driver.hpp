// The Euler time loop of a CMM run on the MI355X engine: initial state, warm-up, steps with the
// reference's time-target clipping, remap logging, monitoring files and saves.
#pragma once

#include "settings.hpp"

namespace cmmhost {

// Cuda-CMM/src/simulation/cmm-simulation-host.cu:1307-1374
double compute_next_timestep(const Settings &s, double t, double dt);

// Cuda-CMM/src/simulation/cmm-euler2d.cu:45-1102 (backward map, fluid particles); returns the exit code
int run_euler(Settings &s);

}  // namespace cmmhost
