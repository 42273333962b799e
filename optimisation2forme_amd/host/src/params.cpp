#include "params.hpp"

#include <fstream>
#include <iostream>
#include <sstream>

namespace navier_stokes {

void SimulationParams::finalize() {
    cv = cp / gamma;
    R = cp - cv;
}

namespace {
std::string trim(const std::string &s) {
    const auto first = s.find_first_not_of(" \t\r\n");
    if (first == std::string::npos) return "";
    const auto last = s.find_last_not_of(" \t\r\n");
    return s.substr(first, last - first + 1);
}
}  // namespace

bool load_config(const std::string &filename, SimulationParams &p) {
    std::ifstream in(filename);
    if (!in) return false;
    std::string line;
    while (std::getline(in, line)) {
        const std::string t = trim(line);
        if (t.empty() || t[0] == '#') continue;
        const auto eq = t.find('=');
        if (eq == std::string::npos) continue;
        const std::string key = trim(t.substr(0, eq));
        const std::string val = trim(t.substr(eq + 1));
        std::istringstream ss(val);
        if (key == "mesh_name") p.mesh_name = val;
        else if (key == "Lx") ss >> p.Lx;
        else if (key == "Ly") ss >> p.Ly;
        else if (key == "Nx") ss >> p.Nx;
        else if (key == "Ny") ss >> p.Ny;
        else if (key == "cfl_number") ss >> p.cfl_number;
        else if (key == "final_time") ss >> p.final_time;
        else if (key == "gamma") ss >> p.gamma;
        else if (key == "viscosity") ss >> p.viscosity;
        else if (key == "Pr") ss >> p.Pr;
        else if (key == "cp") ss >> p.cp;
        else if (key == "cv") ss >> p.cv;
        else if (key == "R") ss >> p.R;
        else if (key == "advection_scheme") p.advection_scheme = val;
        else if (key == "boundary_conditions") p.boundary_conditions = val;
        else if (key == "initial_condition") p.initial_condition = val;
        else if (key == "noise_level") ss >> p.noise_level;
        else if (key == "obstacle_type") p.obstacle_type = val;
        else if (key == "obstacle_center_x") ss >> p.obstacle_center_x;
        else if (key == "obstacle_center_y") ss >> p.obstacle_center_y;
        else if (key == "obstacle_size") ss >> p.obstacle_size;
        else if (key == "inflow_velocity") ss >> p.inflow_velocity;
        else if (key == "rho_ref") ss >> p.rho_ref;
        else if (key == "p_ref") ss >> p.p_ref;
        else if (key == "output_directory") p.output_directory = val;
        else if (key == "save_interval") ss >> p.save_interval;
        else if (key == "state_save_path") p.state_save_path = val;
        else if (key == "state_load_path") p.state_load_path = val;
        else if (key == "output_quantity") p.output_quantity = val;
        else if (key == "verbosity") ss >> p.verbosity;
        else if (key == "reload_config") ss >> p.reload_config;
        else if (key == "shape_opt_enable") ss >> p.shape_opt_enable;
        else if (key == "shape_param_file") p.shape_param_file = val;
        else if (key == "shape_log_file") p.shape_log_file = val;
        else if (key == "limiter") p.limiter = val;
        else if (key == "venkat_K") ss >> p.venkat_K;
        else if (key == "interp_mode") p.interp_mode = val;
        else if (key == "heat_flux") p.heat_flux = val;
        else if (key == "adjoint_tol") ss >> p.adjoint_tol;
        else if (key == "adjoint_eps") ss >> p.adjoint_eps;
        else if (key == "adjoint_chunk") ss >> p.adjoint_chunk;
        else if (key == "adjoint_max_iters") ss >> p.adjoint_max_iters;
        else if (key == "max_steps") ss >> p.max_steps;
        else if (key == "adjoint_operator") p.adjoint_operator = val;
        else if (key == "adjoint_rhs") p.adjoint_rhs = val;
        else if (key == "adjoint_preconditioner") p.adjoint_preconditioner = val;
        else if (key == "shape_gradient") p.shape_gradient = val;
    }
    p.finalize();
    return true;
}

void print_params(const SimulationParams &p) {
    std::cout << "=== Simulation parameters ===\n"
              << " Mesh:          " << p.mesh_name << "\n"
              << " Viscosity:     " << p.viscosity << "\n"
              << " Pr:            " << p.Pr << "\n"
              << " Mach:          " << p.inflow_velocity << "\n"
              << " rho_ref:       " << p.rho_ref << " kg/m^3\n"
              << " p_ref:         " << p.p_ref << " Pa\n"
              << " CFL:           " << p.cfl_number << "\n"
              << " BCs:           " << p.boundary_conditions << "\n"
              << " Init:          " << p.initial_condition << "\n"
              << " Limiter:       " << p.limiter << " (K=" << p.venkat_K << ")\n"
              << " Interpolation: " << p.interp_mode << "\n"
              << " Heat flux:     " << p.heat_flux << "\n"
              << " Save every:    " << p.save_interval << "\n"
              << " Final time:    " << p.final_time << "\n"
              << "=============================\n"
              << std::flush;
}

o2f_params to_o2f_params(const SimulationParams &p) {
    o2f_params o;
    o2f_params_default(&o);
    o.gamma = p.gamma;
    o.viscosity = p.viscosity;
    o.Pr = p.Pr;
    o.cp = p.cp;
    o.cfl_number = p.cfl_number;
    o.rho_ref = p.rho_ref;
    o.p_ref = p.p_ref;
    o.inflow_velocity = p.inflow_velocity;
    o.venkat_k = p.venkat_K;
    o.limiter = (p.limiter == "barth_jespersen" || p.limiter == "bj") ? O2F_LIMITER_BJ : O2F_LIMITER_VENKAT;
    o.interp_mode = p.interp_mode == "mvc" ? O2F_INTERP_MVC : O2F_INTERP_REF;
    o.heat_flux = p.heat_flux == "fourier" ? O2F_HEAT_FLUX_FOURIER : O2F_HEAT_FLUX_REF;
    o.bc_inflow_outflow_x = p.boundary_conditions == "inflow_outflow_x" ? 1 : 0;
    return o;
}

}  // namespace navier_stokes
