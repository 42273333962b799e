// SimulationParams and the config.txt reader of the reference driver
// (include/params.hpp:15-81, src/params.cpp:5-114), plus the keys the gfx950
// path adds.  Parsing semantics are the reference's: `key = value` lines,
// '#' comment lines, surrounding whitespace trimmed, unknown keys ignored,
// the last occurrence of a key wins.
#ifndef O2F_HOST_PARAMS_HPP
#define O2F_HOST_PARAMS_HPP

#include <ostream>
#include <string>

#include "o2f.h"

namespace navier_stokes {

struct SimulationParams {
    // domain and mesh
    std::string mesh_name = "naca_2D.msh";
    double Lx = 1.0, Ly = 1.0;
    int Nx = 128, Ny = 128;
    // physics and time
    double cfl_number = 0.1;
    double final_time = 5.0;
    double gamma = 1.4;
    double viscosity = 1e-6;
    double Pr = 0.71;
    double cp = 1005.0;
    double cv = 718.0;
    double R = 287.0;
    // schemes
    std::string advection_scheme = "compute_advection_weno3_HLLC";
    std::string boundary_conditions = "periodic";
    int bc_ghost_layers = 0;
    // initial condition
    std::string initial_condition = "one";
    double noise_level = 0.0;
    // obstacle (parsed, unused by the hot path)
    std::string obstacle_type = "none";
    double obstacle_center_x = 0.5, obstacle_center_y = 0.5, obstacle_size = 0.1;
    // inflow / outflow
    double inflow_velocity = 1.0;
    double rho_ref = 1.225;
    double p_ref = 101325.0;
    // output
    std::string output_directory = "./results";
    int save_interval = 1;
    std::string state_save_path = "";
    std::string state_load_path = "";
    std::string output_quantity = "schlieren";
    int verbosity = 1;
    int reload_config = 0;
    // shape optimisation
    bool shape_opt_enable = false;
    std::string shape_param_file = "";
    std::string shape_log_file = "";

    // ---- keys added by the gfx950 path (defaults reproduce the reference) ----
    std::string limiter = "venkat";   // "venkat" | "barth_jespersen"
    double venkat_K = 1e-2;           // k_eps of computeVenkatAlpha
    std::string interp_mode = "ref";  // "ref" | "mvc" (valence >= 5 vertices)
    std::string heat_flux = "reference";  // "reference" (R x Fourier) | "fourier" (extension)
    double adjoint_tol = 1e-6;        // BiCGSTAB tolerance
    double adjoint_eps = 1e-6;        // FD step of the flux blocks and the shape partials
    int adjoint_chunk = 500;          // BiCGSTAB restart period
    int adjoint_max_iters = 100000;   // BiCGSTAB iteration cap
    int max_steps = -1;               // cap on time steps (-1: none)
    // adjoint and shape gradient of optimizeShape (defaults: the reference's)
    std::string adjoint_operator = "reference";   // "reference" (FD HLLC blocks) | "exact"
    std::string adjoint_rhs = "reference";        // "reference" | "exact" (with the exact operator)
    std::string adjoint_preconditioner = "none";  // "none" | "block_jacobi"
    std::string shape_gradient = "finite_difference";  // | "tangent" | "tangent_full" (forward mode)
    // 1: steps between two host events run as one device-gated batch; 0: the
    // reference's loop body on the host, one dt reduction and one
    // step_SSP_RK3 (three rk3_substep calls) per step
    int device_time_loop = 1;
    // "none" | "rcm" | "rcm_xcd": renumber the mesh for the device's gathers
    // (o2f_mesh_reorder); state files keep the mesh file's cell order
    std::string mesh_reorder = "none";
    // stages of the time loop (o2f_set_tile_cells): -1 fused tile stages
    // sized automatically, 0 the four-kernel stages, n > 0 tiles of at most
    // n cells (results bitwise the same)
    int tile_cells = -1;

    // c_v = c_p / gamma, R = c_p - c_v (src/params.cpp:5-8)
    void finalize();
};

// false when the file cannot be opened (src/params.cpp:18-83)
bool load_config(const std::string &filename, SimulationParams &p);
void print_params(const SimulationParams &p);
// every key with its value, `key=value` one per line (navier_stokes --print-config)
void dump_config(const SimulationParams &p, std::ostream &out);
// The members of SimulationParams the device path reads.
o2f_params to_o2f_params(const SimulationParams &p);

}  // namespace navier_stokes

#endif
