// The reference's simulation program (src/app/main.cu: run_simulation on the
// config path given on the command line) over the library: used by the GPU
// end-to-end test.
#include <exception>
#include <iostream>

#include "lsmps3d_amd.hpp"

int main(int argc, char** argv) {
  if (argc > 2) {
    lsmps3d::print_simulation_usage(std::cerr, argv[0]);
    return 2;
  }
  try {
    return lsmps3d::run_simulation(argc == 2 ? argv[1] : "config/simulation.ini");
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
}
