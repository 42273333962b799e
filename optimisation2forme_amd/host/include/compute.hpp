// Wall forces (include/compute.hpp:18) on the gfx950 path.
#ifndef O2F_HOST_COMPUTE_HPP
#define O2F_HOST_COMPUTE_HPP

#include "mesh.hpp"
#include "params.hpp"

namespace navier_stokes {

// C_L, C_D from the face states of the last residual (src/compute.cpp:92-132).
void computeLiftDragCoefficients(const Mesh &mesh, const SimulationParams &params, double &C_L, double &C_D);

}  // namespace navier_stokes

#endif
