// Initial conditions (include/init.hpp:16): the freestream cold start.
#ifndef O2F_HOST_INIT_HPP
#define O2F_HOST_INIT_HPP

#include "mesh.hpp"
#include "params.hpp"

namespace navier_stokes {

// "one_x": uniform freestream along +x (src/init.cpp:59-81, :376-391).  The
// other initial conditions of the reference are test fields outside the
// hot path and raise std::invalid_argument here.
void initialize_flow_field(Mesh &mesh, const SimulationParams &params);

}  // namespace navier_stokes

#endif
