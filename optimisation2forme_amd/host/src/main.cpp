// navier_stokes [config_file]: the reference driver (src/main.cpp:33-94) on
// the gfx950 path.  The SFML window and pixel map are not part of it.
#include <filesystem>
#include <iostream>

#include "mesh.hpp"
#include "params.hpp"
#include "simulation.hpp"

int main(int argc, char *argv[]) {
    using namespace navier_stokes;
    std::string cfg_file = "config.txt";
    bool dump = false;
    for (int i = 1; i < argc; ++i) {
        const std::string arg = argv[i];
        if (arg == "-h" || arg == "--help") {
            std::cout << "Usage: " << argv[0] << " [--print-config] [config_file]\n"
                      << "If no config_file is given, uses \"config.txt\" in CWD.\n";
            return 0;
        }
        if (arg == "--print-config") dump = true;
        else cfg_file = arg;
    }
    SimulationParams params;
    if (!load_config(cfg_file, params)) {
        std::cerr << "Error: impossible de lire le fichier de config '" << cfg_file << "'\n";
        return 1;
    }
    params.finalize();
    if (dump) {  // parsed values, one per line (used by the tests)
        std::cout.precision(17);
        std::cout << "mesh_name=" << params.mesh_name << "\ncfl_number=" << params.cfl_number
                  << "\nfinal_time=" << params.final_time << "\nviscosity=" << params.viscosity
                  << "\nPr=" << params.Pr << "\ninflow_velocity=" << params.inflow_velocity
                  << "\nsave_interval=" << params.save_interval << "\nreload_config=" << params.reload_config
                  << "\nstate_load_path=" << params.state_load_path << "\nstate_save_path=" << params.state_save_path
                  << "\nshape_opt_enable=" << params.shape_opt_enable << "\nshape_param_file=" << params.shape_param_file
                  << "\nlimiter=" << params.limiter << "\nvenkat_K=" << params.venkat_K
                  << "\ninterp_mode=" << params.interp_mode << "\nheat_flux=" << params.heat_flux
                  << "\nadjoint_operator=" << params.adjoint_operator << "\nadjoint_rhs=" << params.adjoint_rhs
                  << "\nadjoint_preconditioner=" << params.adjoint_preconditioner
                  << "\nshape_gradient=" << params.shape_gradient << "\ncv=" << params.cv
                  << "\nR=" << params.R << "\n";
        return 0;
    }
    if (params.verbosity > 0) print_params(params);
    try {
        Mesh mesh(params);
        std::filesystem::create_directories(params.output_directory);
        Simulator simulation(mesh, params);
        simulation.run();
    } catch (const std::exception &e) {
        std::cerr << "Simulation failed: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
