// Velocity + pressure manufactured-solution run through the reference's
// interface (the calling pattern of test/full_test.cpp): Ethier-Steinman on
// [0,1]x[0,1]x[-1,1], walls on every face, N nodes per axis.
// Usage: full_test N steps [final_time Re [mixed]]
// Prints: L1 L2 Linf of velocity, pressure and pressure gradient.
#include <cstdio>
#include <stdexcept>
#include <cstdlib>
#include <iostream>

#include "Manufactured.h"
#include "Norms.h"
#include "PressureEquation.h"
#include "PressureGradient.h"
#include "StaggeredTensorMacros.h"
#include "Timestep.h"

double Reynolds;

static int run(int argc, char *argv[]) {
  using namespace mif;
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s N steps [final_time Re [mixed]]\n", argv[0]);
    return 2;
  }
  const size_t N = std::atol(argv[1]);
  const unsigned steps = std::atoi(argv[2]);
  const Real final_time = argc > 3 ? std::atof(argv[3]) : 1e-4;
  const Real Re = argc > 4 ? std::atof(argv[4]) : 1e3;
  const int precision = (argc > 5 && std::atoi(argv[5])) ? MIF_PRECISION_MIXED : MIF_PRECISION_FP64;
  const std::array<bool, 3> periodic_bc{false, false, false};

  const Constants constants(N, N, N, 1.0, 1.0, 2.0, 0.0, 0.0, -1.0, Re, final_time, steps, 1, 1, 0, periodic_bc,
                            precision);
  PressureSolverStructures structures(constants);
  Reynolds = Re;

  VelocityTensor velocity(constants);
  VelocityTensor velocity_buffer(constants);
  VelocityTensor velocity_buffer_2(constants);
  StaggeredTensor pressure(constants, StaggeringDirection::none);
  StaggeredTensor pressure_buffer(constants, StaggeringDirection::none);
  PressureTensor pressure_solver_buffer(structures);

  TimeVectorFunction exact_velocity(u_exact, v_exact, w_exact);
  velocity.set(exact_velocity.set_time(0.0), true);
  pressure.set([](Real x, Real y, Real z) { return p_exact(0.0, x, y, z); }, true);
  TimeVectorFunction exact_pressure_gradient(dp_dx_exact, dp_dy_exact, dp_dz_exact);

  for (unsigned s = 0; s < steps; s++)
    timestep(velocity, velocity_buffer, velocity_buffer_2, exact_velocity, s * constants.dt, pressure,
             pressure_buffer, pressure_solver_buffer);

  // pressure gradient on the interior through the host mirror, as the
  // reference's driver does, then its boundary values
  VelocityTensor grad(constants);
  VELOCITY_TENSOR_SET_FOR_ALL_POINTS(grad, pressure_gradient_u, pressure_gradient_v, pressure_gradient_w, false,
                                     pressure, i, j, k)
  grad.apply_bc(exact_pressure_gradient.set_time(final_time));

  adjust_pressure(pressure, [final_time](Real x, Real y, Real z) { return p_exact(final_time, x, y, z); });

  const Real e[9] = {
      accumulate_error_mpi_l1(ErrorL1Norm(velocity, exact_velocity, final_time), constants),
      accumulate_error_mpi_l2(ErrorL2Norm(velocity, exact_velocity, final_time), constants),
      accumulate_error_mpi_linf(ErrorLInfNorm(velocity, exact_velocity, final_time), constants),
      accumulate_error_mpi_l1(ErrorL1Norm(pressure, p_exact, final_time), constants),
      accumulate_error_mpi_l2(ErrorL2Norm(pressure, p_exact, final_time), constants),
      accumulate_error_mpi_linf(ErrorLInfNorm(pressure, p_exact, final_time), constants),
      accumulate_error_mpi_l1(ErrorL1Norm(grad, exact_pressure_gradient, final_time), constants),
      accumulate_error_mpi_l2(ErrorL2Norm(grad, exact_pressure_gradient, final_time), constants),
      accumulate_error_mpi_linf(ErrorLInfNorm(grad, exact_pressure_gradient, final_time), constants),
  };
  for (int q = 0; q < 9; q++) std::printf("%.17g%c", e[q], q == 8 ? '\n' : ' ');
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
