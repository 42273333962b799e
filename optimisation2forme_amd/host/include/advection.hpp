// Convective operator (include/advection.hpp:16-19) on the gfx950 path.
#ifndef O2F_HOST_ADVECTION_HPP
#define O2F_HOST_ADVECTION_HPP

#include "boundary.hpp"
#include "mesh.hpp"
#include "params.hpp"

namespace navier_stokes {

// HLLC flux of one edge from its stored face states (src/advection.cpp:335-412).
void compute_hllc_flux(Edge &edge, const SimulationParams &params, double &edge_flux_rho, double &edge_flux_rho_u,
                       double &edge_flux_rho_v, double &edge_flux_E);
// Cell gradients, limited face states, HLLC, R (src/advection.cpp:415-466).
void computeAdvectionTerm(Mesh &mesh, const SimulationParams &params);

}  // namespace navier_stokes

#endif
