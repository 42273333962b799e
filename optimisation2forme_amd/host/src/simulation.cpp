// Simulator::run / time_loop of the reference (src/simulation.cpp:40-232),
// device resident: the state goes to the GPU once, steps run there, and the
// host only wakes at output / reload steps.
#include "simulation.hpp"

#include <algorithm>
#include <iostream>

#include "compute.hpp"
#include "init.hpp"
#include "shape_optimization.hpp"
#include "utils.hpp"

namespace navier_stokes {

Simulator::Simulator(Mesh &mesh, const SimulationParams &params) : mesh_(mesh), params_(params) {}

void Simulator::run() {
    initialize_fields();
    time_loop();
    saveCellStates(mesh_, params_);
    if (params_.shape_opt_enable) optimizeShape(mesh_, params_, params_.adjoint_eps);
}

void Simulator::initialize_fields() {
    initialize_flow_field(mesh_, params_);
    loadCellStates(mesh_, params_);
}

// while (t < final_time) { dt; step; output every save_interval; reload every
// reload_config; t += dt; ++it; }  (src/simulation.cpp:68-126).  Steps between
// two host events run as one device-gated batch.
void Simulator::time_loop() {
    if (!params_.device_time_loop) {
        host_time_loop();
        return;
    }
    mesh_.push_state();
    double t = 0.0;
    int it = 0;
    for (;;) {
        if (params_.max_steps >= 0 && it >= params_.max_steps) break;
        // next step index after which the host has something to do
        int next = it;
        const int si = std::max(params_.save_interval, 1);
        int to_save = (si - it % si) % si;
        int to_reload = params_.reload_config != 0
                            ? (params_.reload_config - it % params_.reload_config) % params_.reload_config
                            : 1 << 30;
        next = it + std::min(to_save, to_reload);
        int n = next - it + 1;
        if (params_.max_steps >= 0) n = std::min(n, params_.max_steps - it);
        int steps = 0;
        double t_start = 0.0, dt = 0.0;
        mesh_.set_params(params_);
        mesh_.check(o2f_run_until(mesh_.ctx(), params_.final_time, n, &steps, &t, &t_start, &dt));
        it += steps;
        if (steps < n) break;  // t reached final_time
        const int last = it - 1;
        if (last % si == 0) output_step(last, t_start, dt);
        // reload_config == 0 means never (the reference evaluates it % 0 first)
        if (params_.reload_config != 0 && last % params_.reload_config == 0) load_config("config.txt", params_);
    }
    mesh_.pull_state();
    mesh_.pull_faces();
}

// The loop body of src/simulation.cpp:75-124 on the host, one step at a
// time: the dt reduction, step_SSP_RK3, output and reload.
void Simulator::host_time_loop() {
    mesh_.push_state();
    double t = 0.0;
    int it = 0;
    const int si = std::max(params_.save_interval, 1);
    while (t < params_.final_time) {
        if (params_.max_steps >= 0 && it >= params_.max_steps) break;
        mesh_.set_params(params_);
        double dt = 0.0;
        mesh_.check(o2f_compute_dt(mesh_.ctx(), &dt));
        step_SSP_RK3(dt);
        if (it % si == 0) output_step(it, t, dt);
        if (params_.reload_config != 0 && it % params_.reload_config == 0) load_config("config.txt", params_);
        t += dt;
        ++it;
    }
    mesh_.pull_faces();
}

void Simulator::step_SSP_RK3(double dt) {
    mesh_.push_state();
    mesh_.check(o2f_rk3_save_old(mesh_.ctx()));
    rk3_substep(0.0, 1.0, dt);
    rk3_substep(3.0 / 4.0, 1.0 / 4.0, dt);
    rk3_substep(1.0 / 3.0, 2.0 / 3.0, dt);
    mesh_.pull_state();
}

void Simulator::rk3_substep(double c0, double c1, double dt) {
    mesh_.check(o2f_rk3_substep(mesh_.ctx(), c0, c1, dt));
}

void Simulator::output_step(int it, double t, double dt) {
    if (params_.verbosity > 0) std::cout << "it=" << it << "  t=" << t << "  dt=" << dt << '\n' << std::flush;
    double C_L, C_D;
    computeLiftDragCoefficients(mesh_, params_, C_L, C_D);
    std::cout << "C_L=" << C_L << " ; C_D=" << C_D << '\n' << std::flush;
}

}  // namespace navier_stokes
