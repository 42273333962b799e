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
    void output_step(int iter, double t, double dt);
    // the three Shu-Osher stages of src/simulation.cpp:130-171, fused on the device
    void step_SSP_RK3(double dt);
};

}  // namespace navier_stokes

#endif
