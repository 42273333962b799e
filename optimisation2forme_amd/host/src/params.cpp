// config.txt reader.  The file format is the reference driver's
// (src/params.cpp:18-83), because scripts write and rewrite it
// (run_opti.sh:160-190): one `key = value` assignment per line; blanks
// (space, tab, CR, LF) around the key and the value do not count; a line
// whose first non-blank character is '#', or that has no '=', is skipped;
// keys this build does not know are ignored; a key assigned twice keeps its
// last value.  Numbers are read the way `std::istream >>` reads them (the
// longest valid prefix; a value that does not start with a number reads
// as 0).
//
// Design: every key is one row of a table that binds its name to a member of
// SimulationParams.  The same table parses a file and prints the parsed
// values (`navier_stokes --print-config`), so a key added here is known to
// both.
#include "params.hpp"

#include <algorithm>
#include <fstream>
#include <iterator>
#include <iostream>
#include <sstream>
#include <string_view>
#include <variant>

namespace navier_stokes {

void SimulationParams::finalize() {
    cv = cp / gamma;
    R = cp - cv;
}

namespace {

using P = SimulationParams;
using Member = std::variant<std::string P::*, double P::*, int P::*, bool P::*>;

struct Key {
    const char *name;
    Member member;
};

// Sorted by name (byte order), so a key is found by binary search.
constexpr Key kKeys[] = {
    {"Lx", &P::Lx},
    {"Ly", &P::Ly},
    {"Nx", &P::Nx},
    {"Ny", &P::Ny},
    {"Pr", &P::Pr},
    {"R", &P::R},
    {"adjoint_chunk", &P::adjoint_chunk},
    {"adjoint_eps", &P::adjoint_eps},
    {"adjoint_max_iters", &P::adjoint_max_iters},
    {"adjoint_operator", &P::adjoint_operator},
    {"adjoint_preconditioner", &P::adjoint_preconditioner},
    {"adjoint_rhs", &P::adjoint_rhs},
    {"adjoint_tol", &P::adjoint_tol},
    {"advection_scheme", &P::advection_scheme},
    {"boundary_conditions", &P::boundary_conditions},
    {"cfl_number", &P::cfl_number},
    {"cp", &P::cp},
    {"cv", &P::cv},
    {"device_time_loop", &P::device_time_loop},
    {"final_time", &P::final_time},
    {"gamma", &P::gamma},
    {"heat_flux", &P::heat_flux},
    {"inflow_velocity", &P::inflow_velocity},
    {"initial_condition", &P::initial_condition},
    {"interp_mode", &P::interp_mode},
    {"limiter", &P::limiter},
    {"max_steps", &P::max_steps},
    {"mesh_name", &P::mesh_name},
    {"mesh_reorder", &P::mesh_reorder},
    {"noise_level", &P::noise_level},
    {"obstacle_center_x", &P::obstacle_center_x},
    {"obstacle_center_y", &P::obstacle_center_y},
    {"obstacle_size", &P::obstacle_size},
    {"obstacle_type", &P::obstacle_type},
    {"output_directory", &P::output_directory},
    {"output_quantity", &P::output_quantity},
    {"p_ref", &P::p_ref},
    {"reload_config", &P::reload_config},
    {"rho_ref", &P::rho_ref},
    {"save_interval", &P::save_interval},
    {"shape_gradient", &P::shape_gradient},
    {"shape_log_file", &P::shape_log_file},
    {"shape_opt_enable", &P::shape_opt_enable},
    {"shape_param_file", &P::shape_param_file},
    {"state_load_path", &P::state_load_path},
    {"state_save_path", &P::state_save_path},
    {"tile_cells", &P::tile_cells},
    {"venkat_K", &P::venkat_K},
    {"verbosity", &P::verbosity},
    {"viscosity", &P::viscosity},
};

constexpr bool sorted_by_name() {
    for (size_t i = 1; i < std::size(kKeys); ++i)
        if (!(std::string_view(kKeys[i - 1].name) < std::string_view(kKeys[i].name))) return false;
    return true;
}
static_assert(sorted_by_name(), "kKeys must stay sorted by name");

const Key *find_key(std::string_view name) {
    const Key *end = std::end(kKeys);
    const Key *k = std::lower_bound(std::begin(kKeys), end, name,
                                    [](const Key &a, std::string_view n) { return std::string_view(a.name) < n; });
    return (k != end && name == k->name) ? k : nullptr;
}

constexpr std::string_view kBlanks = " \t\r\n";

std::string_view strip(std::string_view s) {
    const size_t a = s.find_first_not_of(kBlanks);
    if (a == std::string_view::npos) return {};
    return s.substr(a, s.find_last_not_of(kBlanks) + 1 - a);
}

// operator>> semantics: the value read, 0 when nothing could be read.
template <class T>
T number(std::string_view text) {
    T v{};
    std::istringstream in{std::string(text)};
    in >> v;
    return v;
}

void assign(P &p, const Key &k, std::string_view value) {
    std::visit(
        [&](auto member) {
            using T = std::remove_reference_t<decltype(p.*member)>;
            if constexpr (std::is_same_v<T, std::string>)
                p.*member = std::string(value);
            else
                p.*member = number<T>(value);
        },
        k.member);
}

// Splits one line into (key, value); false for blank, comment and
// assignment-free lines.
bool split_assignment(std::string_view line, std::string_view &key, std::string_view &value) {
    line = strip(line);
    if (line.empty() || line.front() == '#') return false;
    const size_t eq = line.find('=');
    if (eq == std::string_view::npos) return false;
    key = strip(line.substr(0, eq));
    value = strip(line.substr(eq + 1));
    return true;
}

}  // namespace

bool load_config(const std::string &filename, SimulationParams &p) {
    std::ifstream in(filename);
    if (!in) return false;
    for (std::string line; std::getline(in, line);) {
        std::string_view key, value;
        if (!split_assignment(line, key, value)) continue;
        if (const Key *k = find_key(key)) assign(p, *k, value);
    }
    p.finalize();
    return true;
}

void dump_config(const SimulationParams &p, std::ostream &out) {
    const auto old = out.precision(17);
    for (const Key &k : kKeys) {
        out << k.name << '=';
        std::visit([&](auto member) { out << p.*member; }, k.member);
        out << '\n';
    }
    out.precision(old);
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
