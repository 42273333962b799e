// Simulator of the reference driver (include/simulation.hpp:22-47) without
// the SFML window: the time loop runs on the device.
#ifndef O2F_HOST_SIMULATION_HPP
#define O2F_HOST_SIMULATION_HPP

#include "mesh.hpp"
#include "params.hpp"

namespace navier_stokes {

class Simulator {
   public:
    explicit Simulator(Mesh &mesh, const SimulationParams &params);
    void run();

   private:
    Mesh &mesh_;
    SimulationParams params_;

    void initialize_fields();
    void time_loop();
    void host_time_loop();
    void output_step(int iter, double t, double dt);
    // src/simulation.cpp:130-171: U_old <- U, then the three Shu-Osher
    // substeps, each a residual and the update U = c0 U_old + c1 (U + dt R)
    // on the device-resident state
    void step_SSP_RK3(double dt);
    void rk3_substep(double c0, double c1, double dt);
};

}  // namespace navier_stokes

#endif
