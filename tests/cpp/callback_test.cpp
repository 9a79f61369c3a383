// Boundary data from callbacks the device cannot evaluate: the same runs
// through the reference's interface twice, once with the named exact
// solutions (which the shim recognises and hands to the device as
// descriptors) and once with lambdas wrapping them (which it cannot
// recognise, so it evaluates them on the host into face tables).
// Usage: callback_test MODE N steps
//   MODE walls     timestep, Ethier-Steinman on [0,1]x[0,1]x[-1,1], all walls
//   MODE channel   timestep, the same data on a domain periodic in x and z
//   MODE nhn       timestep_nhn, all walls (boundary and Neumann data)
//   MODE velocity  timestep_velocity, the manufactured velocity field
//   MODE apply_bc  VelocityTensor::apply_bc and the non-homogeneous solve alone
// Prints: the largest relative difference between the two runs of u, v, w
// and of the pressure (0 for the velocity-only mode).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "Manufactured.h"
#include "ManufacturedVelocity.h"
#include "PressureEquation.h"
#include "Timestep.h"
#include "TimestepVelocity.h"

double Reynolds;

namespace {

using namespace mif;

double max_rel(const StaggeredTensor &a, const StaggeredTensor &b) {
  const auto &s = a.sizes();
  double num = 0, den = 0;
  for (size_t k = 0; k < s[2]; k++)
    for (size_t j = 0; j < s[1]; j++)
      for (size_t i = 0; i < s[0]; i++) {
        num = std::max(num, std::fabs(a(i, j, k) - b(i, j, k)));
        den = std::max(den, std::fabs(b(i, j, k)));
      }
  return den > 0 ? num / den : num;
}

struct Run {
  std::unique_ptr<VelocityTensor> vel, buf, buf2;
  std::unique_ptr<StaggeredTensor> p, pb;
  explicit Run(const Constants &c)
      : vel(new VelocityTensor(c)),
        buf(new VelocityTensor(c)),
        buf2(new VelocityTensor(c)),
        p(new StaggeredTensor(c, StaggeringDirection::none)),
        pb(new StaggeredTensor(c, StaggeringDirection::none)) {}
};

// Wrap a named function so that the shim cannot recognise it.
TimeFunction hide(double (*f)(double, double, double, double)) {
  const double one = 1.0;
  return [f, one](Real t, Real x, Real y, Real z) { return one * f(t, x, y, z); };
}

int run(int argc, char *argv[]) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s MODE N steps\n", argv[0]);
    return 2;
  }
  const std::string mode = argv[1];
  const size_t N = std::atol(argv[2]);
  const unsigned steps = std::atoi(argv[3]);
  const bool channel = mode == "channel";
  const std::array<bool, 3> periodic_bc{channel, false, channel};
  const Real Re = 1e3;
  const Real T = 1e-3;
  Reynolds = Re;
  const Constants constants(N, N, N, 1.0, 1.0, 2.0, 0.0, 0.0, -1.0, Re, T, steps, 1, 1, 0, periodic_bc);
  PressureSolverStructures structures(constants);
  PressureTensor solver_buffer(structures);

  const bool velocity_only = mode == "velocity";
  const TimeVectorFunction named = velocity_only ? TimeVectorFunction(u_exact_v_test, v_exact_v_test, w_exact_v_test)
                                                 : TimeVectorFunction(u_exact, v_exact, w_exact);
  const TimeVectorFunction hidden = velocity_only ? TimeVectorFunction(hide(u_exact_v_test), hide(v_exact_v_test),
                                                                       hide(w_exact_v_test))
                                                  : TimeVectorFunction(hide(u_exact), hide(v_exact), hide(w_exact));
  const TimeVectorFunction grad_named(dp_dx_exact, dp_dy_exact, dp_dz_exact);
  const TimeVectorFunction grad_hidden(hide(dp_dx_exact), hide(dp_dy_exact), hide(dp_dz_exact));
  if (named.has_descriptor() != true || hidden.has_descriptor() != false || grad_hidden.has_descriptor())
    throw std::runtime_error("the shim recognised a lambda or missed a named function");

  Run a(constants), b(constants);
  for (Run *r : {&a, &b}) {
    r->vel->set(named.set_time(0.0), true);
    r->p->set([](Real x, Real y, Real z) { return p_exact(0.0, x, y, z); }, true);
  }
  for (unsigned s = 0; s < steps; s++) {
    const Real t = s * constants.dt;
    if (mode == "walls" || mode == "channel") {
      timestep(*a.vel, *a.buf, *a.buf2, named, t, *a.p, *a.pb, solver_buffer);
      timestep(*b.vel, *b.buf, *b.buf2, hidden, t, *b.p, *b.pb, solver_buffer);
    } else if (mode == "nhn") {
      timestep_nhn(*a.vel, *a.buf, *a.buf2, named, grad_named, t, *a.p, *a.pb, solver_buffer);
      timestep_nhn(*b.vel, *b.buf, *b.buf2, hidden, grad_hidden, t, *b.p, *b.pb, solver_buffer);
    } else if (velocity_only) {
      timestep_velocity(*a.vel, *a.buf, *a.buf2, named, t);
      timestep_velocity(*b.vel, *b.buf, *b.buf2, hidden, t);
    } else if (mode == "apply_bc") {
      // perturb the interior so that the boundary pass has something to overwrite
      for (Run *r : {&a, &b})
        for (StaggeredTensor *c : r->vel->components) {
          const auto &sz = c->sizes();
          for (size_t q = 0; q < sz[0] * sz[1] * sz[2]; q++) static_cast<Real *>(c->raw_data())[q] += 0.25 * (q % 7);
        }
      a.vel->apply_bc(named.set_time(0.3 + t));
      b.vel->apply_bc(hidden.set_time(0.3 + t));
      solve_pressure_equation_non_homogeneous_neumann(*a.p, solver_buffer, *a.vel,
                                                      grad_named.get_difference_over_time(0.1, 0.4), 0.01);
      solve_pressure_equation_non_homogeneous_neumann(*b.p, solver_buffer, *b.vel,
                                                      grad_hidden.get_difference_over_time(0.1, 0.4), 0.01);
    } else {
      throw std::invalid_argument("unknown mode " + mode);
    }
  }
  double dv = 0;
  for (int c = 0; c < 3; c++) dv = std::max(dv, max_rel(*b.vel->components[c], *a.vel->components[c]));
  const double dp = velocity_only ? 0.0 : max_rel(*b.p, *a.p);
  std::printf("%.17g %.17g\n", dv, dp);
  return 0;
}

}  // namespace

int main(int argc, char *argv[]) {
  try {
    return run(argc, argv);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
