// Velocity-only manufactured run through the reference's interface (the
// calling pattern of test/velocity_test.cpp): unit cube, walls on every
// face, N nodes per axis.
// Usage: velocity_test N steps [final_time Re]
// Prints: L1 L2 Linf of the velocity error.
#include <cstdio>
#include <stdexcept>
#include <cstdlib>

#include "ManufacturedVelocity.h"
#include "Norms.h"
#include "TimestepVelocity.h"

double Reynolds;

static int run(int argc, char *argv[]) {
  using namespace mif;
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s N steps [final_time Re]\n", argv[0]);
    return 2;
  }
  const size_t N = std::atol(argv[1]);
  const unsigned steps = std::atoi(argv[2]);
  const Real final_time = argc > 3 ? std::atof(argv[3]) : 1e-4;
  const Real Re = argc > 4 ? std::atof(argv[4]) : 1e4;
  const std::array<bool, 3> periodic_bc{false, false, false};
  const Constants constants(N, N, N, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, Re, final_time, steps, 1, 1, 0, periodic_bc);
  Reynolds = Re;

  VelocityTensor velocity(constants);
  VelocityTensor velocity_buffer(constants);
  VelocityTensor rhs_buffer(constants);
  TimeVectorFunction exact_velocity(u_exact_v_test, v_exact_v_test, w_exact_v_test);
  velocity.set(exact_velocity.set_time(0.0), true);

  for (unsigned s = 0; s < steps; s++)
    timestep_velocity(velocity, velocity_buffer, rhs_buffer, exact_velocity, s * constants.dt);

  std::printf("%.17g %.17g %.17g\n",
              accumulate_error_mpi_l1(ErrorL1Norm(velocity, exact_velocity, final_time), constants),
              accumulate_error_mpi_l2(ErrorL2Norm(velocity, exact_velocity, final_time), constants),
              accumulate_error_mpi_linf(ErrorLInfNorm(velocity, exact_velocity, final_time), constants));
  return 0;
}

int main(int argc, char *argv[]) {
  try {
    return run(argc, argv);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
