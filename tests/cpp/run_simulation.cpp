// The reference's simulation program (src/app/main.cu: run_simulation on the
// config path given on the command line, usage and failures reported on stderr
// with EXIT_FAILURE) over the library: used by the GPU end-to-end test.
#include <cstdlib>
#include <exception>
#include <iostream>

#include "lsmps3d_amd.hpp"

int main(int argc, char** argv) {
  try {
    if (argc > 2) {
      lsmps3d::print_simulation_usage(std::cerr, argv[0]);
      return EXIT_FAILURE;
    }
    return lsmps3d::run_simulation(argc == 2 ? argv[1] : "config/simulation.ini");
  } catch (const std::exception& e) {
    std::cerr << "Simulation failed: " << e.what() << std::endl;
    return EXIT_FAILURE;
  }
}
