// navier_stokes [config_file]: the reference driver (src/main.cpp:33-94) on
// the gfx950 path.  The SFML window and pixel map are not part of it.
#include <filesystem>
#include <iostream>

#include "mesh.hpp"
#include "params.hpp"
#include "shape_optimization.hpp"
#include "simulation.hpp"

int main(int argc, char *argv[]) {
    using namespace navier_stokes;
    std::string cfg_file = "config.txt";
    bool dump = false;
    for (int i = 1; i < argc; ++i) {
        const std::string arg = argv[i];
        if (arg == "-h" || arg == "--help") {
            std::cout << "Usage: " << argv[0] << " [--print-config] [--print-shape-params file] [config_file]\n"
                      << "If no config_file is given, uses \"config.txt\" in CWD.\n";
            return 0;
        }
        if (arg == "--print-config") {
            dump = true;
        } else if (arg == "--print-shape-params" && i + 1 < argc) {
            // the parsed shape-parameter file: `name=value` in file order,
            // then `optimised=` and the "*=" names (used by the tests)
            std::vector<std::string> names_all, names;
            std::vector<double> values;
            try {
                loadParamsGeneric(argv[++i], names_all, values, names);
            } catch (const std::exception &e) {
                std::cerr << e.what() << "\n";
                return 1;
            }
            std::cout.precision(17);
            for (size_t k = 0; k < names_all.size(); ++k) std::cout << names_all[k] << '=' << values[k] << '\n';
            std::cout << "optimised=";
            for (size_t k = 0; k < names.size(); ++k) std::cout << (k ? "," : "") << names[k];
            std::cout << '\n';
            return 0;
        } else {
            cfg_file = arg;
        }
    }
    SimulationParams params;
    if (!load_config(cfg_file, params)) {
        std::cerr << "Error: impossible de lire le fichier de config '" << cfg_file << "'\n";
        return 1;
    }
    params.finalize();
    if (dump) {  // parsed values, one per line (used by the tests)
        dump_config(params, std::cout);
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
