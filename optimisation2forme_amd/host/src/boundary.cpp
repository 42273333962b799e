// Host versions of the boundary-state overrides (src/boundary.cpp:7-127).
#include "boundary.hpp"

#include <cmath>
#include <stdexcept>

namespace navier_stokes {

namespace {
void inflow_outflow(const SimulationParams &params, char bc, double &rho, double &rho_u, double &rho_v,
                    double &E) {
    const double rho0 = params.rho_ref, p0 = params.p_ref;
    const double a0 = std::sqrt(params.gamma * p0 / rho0);
    const double u0 = params.inflow_velocity * a0, v0 = 0.0;
    const double kinetic = 0.5 * rho0 * (u0 * u0 + v0 * v0);
    const double E0 = params.p_ref / (params.gamma - 1.0) + kinetic;
    if (bc == 'i' || bc == 'f') {  // inlet, far field: freestream
        rho = rho0;
        rho_u = rho0 * u0;
        rho_v = rho0 * v0;
        E = E0;
    } else if (bc == 'w') {  // no-slip wall keeps the internal energy
        const double kinetic_old = 0.5 * (rho_u * rho_u + rho_v * rho_v) / rho;
        const double e_internal = E - kinetic_old;
        rho_u = 0.0;
        rho_v = 0.0;
        E = e_internal;
    }
}
}  // namespace

void apply_boundary_conditions(const SimulationParams &params, const char bc, double &rho, double &rho_u,
                               double &rho_v, double &E) {
    if (params.boundary_conditions != "inflow_outflow_x")
        throw std::invalid_argument("Unknown boundary condition: " + params.boundary_conditions);
    inflow_outflow(params, bc, rho, rho_u, rho_v, E);
}

void apply_diffusion_bc(const SimulationParams &, const char bc, double &tau_xx, double &tau_yy,
                        double &tau_xy_yx, double &q_x, double &q_y) {
    if (bc == 'f' || bc == 'w') tau_xx = tau_yy = tau_xy_yx = q_x = q_y = 0.0;
}

void apply_InterfaceStates_boundary_conditions(const SimulationParams &params, const char bc, double rho_L,
                                               double rho_u_L, double rho_v_L, double E_L, double &rho_R,
                                               double &rho_u_R, double &rho_v_R, double &E_R, double &, double &) {
    rho_R = rho_L;
    rho_u_R = rho_u_L;
    rho_v_R = rho_v_L;
    E_R = E_L;
    if (bc == 'i' || bc == 'o' || bc == 'f') {
        inflow_outflow(params, bc, rho_R, rho_u_R, rho_v_R, E_R);
    } else if (bc == 'w') {  // mirrored momentum ghost
        rho_u_R = -rho_u_L;
        rho_v_R = -rho_v_L;
    }
}

}  // namespace navier_stokes
