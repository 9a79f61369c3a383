// The `mif` driver (src/main.cpp:40-177 of the reference): the lid-driven
// cavity at Re = 1000 read from an input file, run on one GPU, written as
// solution.vtk and profile*.dat.
//
//   mif <input parameter file>
//
// Test case 1: [0,1] x [0,1] x [-1,1], walls everywhere, the lid v = 1 on
// x = 1.  Test case 2: [-0.5,0.5]^3, periodic in z, the lid on x = -0.5.
// The whole domain lives on one GPU, so Py and Pz (processors per axis for
// the reference's MPI decomposition) are checked and then not used.  The
// environment variable MIF_PRECISION=mixed selects the mixed FP32/FP64
// form (default fp64).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <iostream>

#include "InputParser.h"
#include "PressureTensor.h"
#include "TestCaseBoundaries.h"
#include "Timestep.h"
#include "VTKDatExport.h"

double Reynolds;

namespace {

// src/main.cpp:12-37: remove the outputs of an earlier run
void delete_old_files() {
  char filename[256];
  for (int i = 1; i < 4; i++) {
    std::snprintf(filename, sizeof(filename), "profile%d.dat", i);
    std::remove(filename);
  }
  std::remove("solution.vtk");
}

int run(const std::string &input_file) {
  using namespace mif;
  size_t Nx_global = 0, Ny_global = 0, Nz_global = 0;
  Real dt = 0;
  unsigned int num_time_steps = 0;
  int Py = 0, Pz = 0;
  bool test_case_2 = false;
  try {
    parse_input_file(input_file, Nx_global, Ny_global, Nz_global, dt, num_time_steps, Py, Pz, test_case_2);
    if (Pz < 1 || Py < 1) {
      std::cerr << "The number of processors in each direction must be at least 1." << std::endl;
      return 1;
    }
  } catch (const std::exception &ex) {
    std::cerr << "Error parsing input file: " << ex.what() << std::endl;
    return 1;
  }
  if (Py * Pz != 1)
    std::cerr << "note: Py = " << Py << ", Pz = " << Pz << " ignored: one GPU holds the whole domain" << std::endl;

  const char *prec = std::getenv("MIF_PRECISION");
  const int precision = (prec && std::strcmp(prec, "mixed") == 0) ? MIF_PRECISION_MIXED : MIF_PRECISION_FP64;

  constexpr Real Re = 1e3;
  const std::array<bool, 3> periodic_bc{false, false, test_case_2};
  const Constants constants(Nx_global, Ny_global, Nz_global, 1.0, 1.0, test_case_2 ? 1.0 : 2.0,
                            test_case_2 ? -0.5 : 0.0, test_case_2 ? -0.5 : 0.0, test_case_2 ? -0.5 : -1.0, Re,
                            dt * num_time_steps, num_time_steps, 1, 1, 0, periodic_bc, precision);
  PressureSolverStructures structures(constants);
  Reynolds = Re;

  VelocityTensor velocity(constants);
  VelocityTensor velocity_buffer(constants);
  VelocityTensor velocity_buffer_2(constants);
  StaggeredTensor pressure(constants, StaggeringDirection::none);
  StaggeredTensor pressure_buffer(constants, StaggeringDirection::none);
  PressureTensor pressure_solver_buffer(structures);

  TimeVectorFunction exact_velocity(test_case_2 ? exact_u_t2 : exact_u_t1, test_case_2 ? exact_v_t2 : exact_v_t1,
                                    test_case_2 ? exact_w_t2 : exact_w_t1);
  velocity.set(exact_velocity.set_time(0.0), true);
  pressure.set(test_case_2 ? exact_p_initial_t2 : exact_p_initial_t1, true);

  for (unsigned int time_step = 0; time_step < num_time_steps; time_step++) {
    const Real current_time = time_step * constants.dt;
    timestep(velocity, velocity_buffer, velocity_buffer_2, exact_velocity, current_time, pressure, pressure_buffer,
             pressure_solver_buffer);
  }

  delete_old_files();
  writeVTK("solution.vtk", velocity, pressure);
  if (!test_case_2) {
    writeDat("profile1.dat", velocity, pressure, 1, 0.5, 0.5, 0);
    writeDat("profile2.dat", velocity, pressure, 0, 0.5, 0.5, 0);
  } else {
    writeDat("profile1.dat", velocity, pressure, 1, 0, 0, 0);
    writeDat("profile2.dat", velocity, pressure, 0, 0, 0, 0);
    writeDat("profile3.dat", velocity, pressure, 2, 0, 0, 0);
  }
  return 0;
}

}  // namespace

int main(int argc, char *argv[]) {
  if (argc != 2) {
    std::cerr << "Usage: ./mif [input parameter file]" << std::endl;
    return 1;
  }
  try {
    return run(argv[1]);
  } catch (const std::exception &ex) {
    std::cerr << "mif: " << ex.what() << std::endl;
    return 1;
  }
}
