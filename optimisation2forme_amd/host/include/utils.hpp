// Timers and state files of the reference (include/utils.hpp:26-45).
#ifndef O2F_HOST_UTILS_HPP
#define O2F_HOST_UTILS_HPP

#include <chrono>
#include <string>

#include "mesh.hpp"
#include "params.hpp"

namespace navier_stokes {

using Clock = std::chrono::high_resolution_clock;
using TimePoint = Clock::time_point;

TimePoint time_tic();
double time_toc(const TimePoint &t0, const std::string &msg = "", int verbosity = 1);
// text state files: N, then N lines "rho rho_u rho_v E" (src/utils.cpp:29-107)
void loadCellStates(Mesh &mesh, const SimulationParams &params);
void saveCellStates(const Mesh &mesh, const SimulationParams &params);

}  // namespace navier_stokes

#endif
