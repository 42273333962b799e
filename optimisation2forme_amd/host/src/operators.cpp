// Operator entry points of the reference, dispatched to the gfx950 library.
#include <cstdio>
#include <filesystem>
#include <iostream>
#include <random>
#include <stdexcept>
#include <vector>

#include "advection.hpp"
#include "compute.hpp"
#include "diffusion.hpp"
#include "init.hpp"
#include "utils.hpp"

namespace navier_stokes {

void compute_hllc_flux(Edge &edge, const SimulationParams &params, double &f_rho, double &f_rho_u,
                       double &f_rho_v, double &f_E) {
    const double qL[4] = {edge.rho_L, edge.rho_u_L, edge.rho_v_L, edge.E_L};
    const double qR[4] = {edge.rho_R, edge.rho_u_R, edge.rho_v_R, edge.E_R};
    const double n[2] = {edge.edgeNormal.x, edge.edgeNormal.y};
    double F[4];
    if (o2f_hllc_flux_batch(qL, qR, n, params.gamma, F, 1) != O2F_OK) throw std::runtime_error(o2f_global_error());
    f_rho = F[0];
    f_rho_u = F[1];
    f_rho_v = F[2];
    f_E = F[3];
}

void computeAdvectionTerm(Mesh &mesh, const SimulationParams &params) {
    mesh.set_params(params);
    mesh.push_state();
    std::vector<double> R(4 * mesh.cells.size());
    mesh.check(o2f_advection(mesh.ctx(), R.data()));
    mesh.store_residual(R.data());
    mesh.pull_gradients();
    mesh.pull_faces();
}

void computeDiffusionTerm(Mesh &mesh, const SimulationParams &params) {
    mesh.set_params(params);
    mesh.push_state();
    std::vector<double> R(4 * mesh.cells.size());
    mesh.check(o2f_diffusion(mesh.ctx(), R.data()));
    mesh.store_residual(R.data());
    // wall stresses are zeroed by the diffusion BC (src/diffusion.cpp:155-162)
    for (Edge &e : mesh.edges)
        if (e.boundaryCondition == 'w') e.tau_xx = e.tau_xy_yx = e.tau_yy = 0.0;
}

void computeLiftDragCoefficients(const Mesh &mesh, const SimulationParams &, double &C_L, double &C_D) {
    mesh.check(o2f_lift_drag(mesh.ctx(), &C_L, &C_D));
}

TimePoint time_tic() { return Clock::now(); }

double time_toc(const TimePoint &t0, const std::string &msg, int verbosity) {
    const double el = std::chrono::duration<double>(Clock::now() - t0).count();
    if (verbosity >= 2 && !msg.empty()) std::cout << "[Timer] " << msg << ": " << el << " s\n";
    return el;
}

void loadCellStates(Mesh &mesh, const SimulationParams &params) {
    const std::string &path = params.state_load_path;
    if (path.empty() || path == "none") return;
    mesh.check(o2f_load_state_file(mesh.ctx(), path.c_str()));
    if (!mesh.cell_perm().empty()) {  // the file is in the mesh file's cell order
        std::vector<double> U(4 * mesh.cells.size());
        mesh.check(o2f_get_state(mesh.ctx(), U.data()));
        mesh.state_from_file_order(U);
        mesh.check(o2f_set_state(mesh.ctx(), U.data()));
    }
    mesh.pull_state();
}

void saveCellStates(const Mesh &mesh, const SimulationParams &params) {
    const std::string &path = params.state_save_path;
    if (path.empty() || path == "none") return;
    std::filesystem::path p(path);
    if (!p.parent_path().empty()) {
        std::error_code ec;
        if (!std::filesystem::create_directories(p.parent_path(), ec) && ec)
            throw std::runtime_error("Impossible de créer le dossier: " + p.parent_path().string());
    }
    if (mesh.cell_perm().empty()) {
        mesh.check(o2f_save_state_file(mesh.ctx(), path.c_str()));
        return;
    }
    // written in the mesh file's cell order, then the state put back
    std::vector<double> U(4 * mesh.cells.size());
    mesh.check(o2f_get_state(mesh.ctx(), U.data()));
    std::vector<double> F(U);
    mesh.state_to_file_order(F);
    mesh.check(o2f_set_state(mesh.ctx(), F.data()));
    const int st = o2f_save_state_file(mesh.ctx(), path.c_str());
    mesh.check(o2f_set_state(mesh.ctx(), U.data()));
    mesh.check(st);
}

void initialize_flow_field(Mesh &mesh, const SimulationParams &params) {
    if (params.initial_condition != "one_x")
        throw std::invalid_argument("Unknown initial condition: " + params.initial_condition);
    mesh.check(o2f_init_freestream(mesh.ctx()));
    mesh.pull_state();
    if (params.noise_level != 0.0) {  // add_velocity_noise, src/init.cpp:362-371
        std::mt19937 gen(std::random_device{}());
        std::uniform_real_distribution<double> d(-params.noise_level, params.noise_level);
        for (Cell &c : mesh.cells) {
            c.rho_u += d(gen);
            c.rho_v += d(gen);
        }
        mesh.push_state();
    }
}

}  // namespace navier_stokes
