// Viscous operator (include/diffusion.hpp:16) on the gfx950 path.
#ifndef O2F_HOST_DIFFUSION_HPP
#define O2F_HOST_DIFFUSION_HPP

#include "mesh.hpp"
#include "params.hpp"

namespace navier_stokes {

// Adds the viscous divergence to R (src/diffusion.cpp:206-243); uses the
// gradients of the last computeAdvectionTerm.
void computeDiffusionTerm(Mesh &mesh, const SimulationParams &params);

}  // namespace navier_stokes

#endif
