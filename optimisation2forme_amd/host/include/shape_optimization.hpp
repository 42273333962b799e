// Discrete adjoint and shape gradient (include/shape_optimization.hpp:24-29)
// on the gfx950 path.
#ifndef O2F_HOST_SHAPE_OPTIMIZATION_HPP
#define O2F_HOST_SHAPE_OPTIMIZATION_HPP

#include <string>
#include <vector>

#include "mesh.hpp"
#include "params.hpp"

namespace navier_stokes {

// computeResidual (src/shape_optimization.cpp:171-189)
void computeResidual(Mesh &mesh, const SimulationParams &params, std::vector<double> &R_rho,
                     std::vector<double> &R_rho_u, std::vector<double> &R_rho_v, std::vector<double> &R_E);
// The "key (=|*=|space) value" shape-parameter file (:19-98).
void loadParamsGeneric(const std::string &filepath, std::vector<std::string> &param_names_all,
                       std::vector<double> &param_values_all, std::vector<std::string> &param_names);
// J, adjoint, shape gradient, preview and log (:1022-1067).
void optimizeShape(Mesh &mesh, const SimulationParams &params, double eps = 1e-6);

}  // namespace navier_stokes

#endif
