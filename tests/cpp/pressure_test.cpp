// Stand-alone pressure solve through the reference's interface (the calling
// pattern of test/pressure_test_hn.cpp and test/pressure_test_nhn.cpp): the
// Poisson test field on [0, 2pi]^3 with N nodes per axis, periodic axes
// given as three 0/1 flags.  Mode hn solves with homogeneous Neumann /
// periodic conditions, nhn with the exact pressure gradient on the walls.
// Usage: pressure_test N mode(hn|nhn) px py pz
// Prints: L1 L2 Linf of the pressure error.
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <cstdlib>
#include <cstring>

#include "ManufacturedPressure.h"
#include "Norms.h"
#include "PressureEquation.h"

static int run(int argc, char *argv[]) {
  using namespace mif;
  if (argc < 6) {
    std::fprintf(stderr, "usage: %s N hn|nhn px py pz\n", argv[0]);
    return 2;
  }
  const size_t N = std::atol(argv[1]);
  const bool nhn = std::strcmp(argv[2], "nhn") == 0;
  const std::array<bool, 3> periodic_bc{std::atoi(argv[3]) != 0, std::atoi(argv[4]) != 0, std::atoi(argv[5]) != 0};
  const Real L = 2 * M_PI;
  const Constants constants(N, N, N, L, L, L, 0.0, 0.0, 0.0, 1.0, 1.0, 1, 1, 1, 0, periodic_bc);
  PressureSolverStructures structures(constants);
  PressureTensor buffer(structures);
  StaggeredTensor pressure(constants, StaggeringDirection::none);
  VelocityTensor velocity(constants);

  const Real time = 1.0;
  TimeVectorFunction exact_velocity(u_exact_p_test, v_exact_p_test, w_exact_p_test);
  velocity.set(exact_velocity.set_time(time), true);
  if (nhn) {
    TimeVectorFunction gradient(dp_dx_exact_p_test, dp_dy_exact_p_test, dp_dz_exact_p_test);
    solve_pressure_equation_non_homogeneous_neumann(pressure, buffer, velocity, gradient.set_time(time), 1.0);
  } else {
    solve_pressure_equation_homogeneous_periodic(pressure, buffer, velocity, 1.0);
  }
  adjust_pressure(pressure, [time](Real x, Real y, Real z) { return p_exact_p_test(time, x, y, z); });
  std::printf("%.17g %.17g %.17g\n", ErrorL1Norm(pressure, p_exact_p_test, time),
              ErrorL2Norm(pressure, p_exact_p_test, time), ErrorLInfNorm(pressure, p_exact_p_test, time));
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
