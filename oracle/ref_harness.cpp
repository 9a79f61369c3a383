/*
 * ref_harness.cpp -- C entry points over the reference solver's own
 * FFTW-free sources, compiled from /root/reference by oracle/Makefile (target
 * `ref`) into oracle/_ref/libmif_ref.so.
 *
 * TEST INFRASTRUCTURE ONLY.  The tests use it to check the oracle's
 * restatement against the reference itself, on the parts of the hot path that
 * build without FFTW and without generated code:
 *   src/Constants.cpp           extents, bases and derived constants
 *   src/StaggeredTensor.cpp     set, apply_periodic_bc
 *   src/VelocityTensor.cpp      set, apply_bc
 *   src/VectorFunction.cpp      set_time, get_difference_over_time
 *   src/Norms.cpp               ErrorL1/L2/LInfNorm
 *   include/MomentumEquation.h  momentum right-hand side (OPT_CPU_CACHE, as
 *                               the reference Makefile:29 builds it)
 *   include/PressureGradient.h, include/VelocityDivergence.h
 *   include/TestCaseBoundaries.h lid-driven boundary data
 *   include/Tensor.h            text dump format
 * Everything runs at one rank (Py = Pz = 1); at one rank none of these calls
 * MPI except the norms' MPI_Allreduce, for which the harness initialises MPI
 * once.
 *
 * Analytic functions come in as a C function pointer plus a context pointer
 * (the oracle's orc_fn_eval and a mif_function descriptor), so both sides
 * evaluate bit-identical function values and the comparison isolates the
 * code under test.
 */
#include "Constants.h"
#include "MomentumEquation.h"
#include "Norms.h"
#include "PressureGradient.h"
#include "StaggeredTensorMacros.h"
#include "TestCaseBoundaries.h"
#include "VectorFunction.h"
#include "VelocityDivergence.h"
#include "VelocityTensor.h"

#include <mpi.h>

#include <array>
#include <cstring>
#include <memory>

#include "../include/mif_types.h"

using namespace mif;

extern "C" {
typedef double (*ref_eval_fn)(const void *ctx, int comp, double t, double x, double y, double z);
}

namespace {

std::unique_ptr<Constants> make_constants(const mif_grid *g) {
  const std::array<bool, 3> per = {g->periodic[0] != 0, g->periodic[1] != 0, g->periodic[2] != 0};
  return std::make_unique<Constants>(g->Nx_global, g->Ny_global, g->Nz_global, g->x_size, g->y_size, g->z_size,
                                     g->min_x, g->min_y, g->min_z, g->Re, g->final_time, g->num_time_steps, 1, 1,
                                     0, per);
}

TimeVectorFunction time_function(ref_eval_fn fn, const void *ctx) {
  return TimeVectorFunction([fn, ctx](Real t, Real x, Real y, Real z) { return fn(ctx, 0, t, x, y, z); },
                            [fn, ctx](Real t, Real x, Real y, Real z) { return fn(ctx, 1, t, x, y, z); },
                            [fn, ctx](Real t, Real x, Real y, Real z) { return fn(ctx, 2, t, x, y, z); });
}

size_t count(const StaggeredTensor &t) {
  const auto &s = t.sizes();
  return s[0] * s[1] * s[2];
}

void load_into(StaggeredTensor &t, const double *src) {
  std::memcpy(t.raw_data(), src, count(t) * sizeof(double));
}

void store_from(StaggeredTensor &t, double *dst) {
  std::memcpy(dst, t.raw_data(), count(t) * sizeof(double));
}

StaggeredTensor &component(VelocityTensor &vel, int stag) {
  return *vel.components[stag];
}

void ensure_mpi() {
  int flag = 0;
  MPI_Initialized(&flag);
  if (!flag) MPI_Init(nullptr, nullptr);
}

}  // namespace

extern "C" {

/* Constants::Constants at one rank (src/Constants.cpp:58-120).
 * ext[4][3]: extents of u, v, w (StaggeredTensor of staggering x, y, z) and of
 * the pressure (staggering none), from the tensors' own sizes().
 * ints[7]: base_i, base_j, base_k, Ny_owner, Nz_owner, prev/next proc (4
 * values packed: prev_y, next_y, prev_z, next_z) -> 9 values.
 * reals[22]: dt, dx, dy, dz, 1/(2h) x3, 1/(8h) x3, 1/(Re h^2) x3, h/2 x3,
 * 1/h x3, and Nx_domains, Ny_domains_global, Nz_domains_global as doubles. */
int ref_constants(const mif_grid *g, long ext[4][3], long ints[9], double reals[22]) {
  const auto c = make_constants(g);
  VelocityTensor vel(*c);
  StaggeredTensor p(*c, StaggeringDirection::none);
  for (int s = 0; s < 3; s++)
    for (int a = 0; a < 3; a++) ext[s][a] = static_cast<long>(vel.components[s]->sizes()[a]);
  for (int a = 0; a < 3; a++) ext[3][a] = static_cast<long>(p.sizes()[a]);
  const long iv[9] = {c->base_i, c->base_j, c->base_k, static_cast<long>(c->Ny_owner),
                      static_cast<long>(c->Nz_owner), c->prev_proc_y, c->next_proc_y, c->prev_proc_z,
                      c->next_proc_z};
  std::memcpy(ints, iv, sizeof(iv));
  const double rv[22] = {c->dt,
                         c->dx, c->dy, c->dz,
                         c->one_over_2_dx, c->one_over_2_dy, c->one_over_2_dz,
                         c->one_over_8_dx, c->one_over_8_dy, c->one_over_8_dz,
                         c->one_over_dx2_Re, c->one_over_dy2_Re, c->one_over_dz2_Re,
                         c->dx_over_2, c->dy_over_2, c->dz_over_2,
                         c->one_over_dx, c->one_over_dy, c->one_over_dz,
                         static_cast<double>(c->Nx_domains), static_cast<double>(c->Ny_domains_global),
                         static_cast<double>(c->Nz_domains_global)};
  std::memcpy(reals, rv, sizeof(rv));
  return 0;
}

/* StaggeredTensor::set / VelocityTensor::set (src/StaggeredTensor.cpp:205-219,
 * src/VelocityTensor.cpp:28-34) of component `comp` of fn at time t, on a
 * field of staggering stag (3 = pressure node). */
int ref_set(const mif_grid *g, int stag, double *f, ref_eval_fn fn, const void *ctx, int comp, double t,
            int include_border) {
  const auto c = make_constants(g);
  const std::function<Real(Real, Real, Real)> fx = [fn, ctx, comp, t](Real x, Real y, Real z) {
    return fn(ctx, comp, t, x, y, z);
  };
  if (stag == 3) {
    StaggeredTensor p(*c, StaggeringDirection::none);
    load_into(p, f);
    p.set(fx, include_border != 0);
    store_from(p, f);
  } else {
    VelocityTensor vel(*c);
    StaggeredTensor &t3 = component(vel, stag);
    load_into(t3, f);
    t3.set(fx, include_border != 0);
    store_from(t3, f);
  }
  return 0;
}

/* velocity.apply_bc(exact_velocity.set_time(t)) as timestep() calls it
 * (src/Timestep.cpp:109, src/VelocityTensor.cpp:36-233). */
int ref_apply_bc(const mif_grid *g, double *u, double *v, double *w, ref_eval_fn fn, const void *ctx, double t) {
  const auto c = make_constants(g);
  VelocityTensor vel(*c);
  load_into(vel.u, u);
  load_into(vel.v, v);
  load_into(vel.w, w);
  const TimeVectorFunction ev = time_function(fn, ctx);
  vel.apply_bc(ev.set_time(t));
  store_from(vel.u, u);
  store_from(vel.v, v);
  store_from(vel.w, w);
  return 0;
}

/* StaggeredTensor::apply_periodic_bc (src/StaggeredTensor.cpp:221-256). */
int ref_apply_periodic(const mif_grid *g, int stag, double *f) {
  const auto c = make_constants(g);
  if (stag == 3) {
    StaggeredTensor p(*c, StaggeringDirection::none);
    load_into(p, f);
    p.apply_periodic_bc();
    store_from(p, f);
  } else {
    VelocityTensor vel(*c);
    StaggeredTensor &t3 = component(vel, stag);
    load_into(t3, f);
    t3.apply_periodic_bc();
    store_from(t3, f);
  }
  return 0;
}

/* calculate_momentum_rhs_{u,v,w} (include/MomentumEquation.h:43-255) at every
 * interior point of component comp, the loop of COMPUTE_COMPONENT_Y2
 * (src/Timestep.cpp:10-22); other points of out are left untouched. */
int ref_momentum_rhs(const mif_grid *g, const double *u, const double *v, const double *w, int comp,
                     double *out) {
  const auto c = make_constants(g);
  VelocityTensor vel(*c);
  load_into(vel.u, u);
  load_into(vel.v, v);
  load_into(vel.w, w);
  StaggeredTensor &t3 = component(vel, comp);
  const auto &s = t3.sizes();
  for (size_t k = 1; k < s[2] - 1; k++)
    for (size_t j = 1; j < s[1] - 1; j++)
      for (size_t i = 1; i < s[0] - 1; i++) {
        const Real r = comp == 0 ? calculate_momentum_rhs_u(vel, i, j, k)
                                 : (comp == 1 ? calculate_momentum_rhs_v(vel, i, j, k)
                                              : calculate_momentum_rhs_w(vel, i, j, k));
        out[i + s[0] * (j + s[1] * k)] = r;
      }
  return 0;
}

/* pressure_gradient_{u,v,w} (include/PressureGradient.h:9-25) at every
 * interior point of component comp (the loop of
 * update_velocity_pressure_gradient, src/Timestep.cpp:66-72). */
int ref_pressure_gradient(const mif_grid *g, const double *p, int comp, double *out) {
  const auto c = make_constants(g);
  StaggeredTensor pt(*c, StaggeringDirection::none);
  load_into(pt, p);
  VelocityTensor vel(*c);
  const auto &s = vel.components[comp]->sizes();
  for (size_t k = 1; k < s[2] - 1; k++)
    for (size_t j = 1; j < s[1] - 1; j++)
      for (size_t i = 1; i < s[0] - 1; i++) {
        const Real r = comp == 0 ? pressure_gradient_u(pt, i, j, k)
                                 : (comp == 1 ? pressure_gradient_v(pt, i, j, k) : pressure_gradient_w(pt, i, j, k));
        out[i + s[0] * (j + s[1] * k)] = r;
      }
  return 0;
}

/* calculate_velocity_divergence (include/VelocityDivergence.h:9-20) / dt over
 * the owner points of the pressure (STAGGERED_TENSOR_ITERATE_OVER_ALL_OWNER_
 * POINTS, include/StaggeredTensorMacros.h:37-84), the loop of
 * compute_rhs_homogeneous_periodic (src/PressureEquation.cpp:59-61). */
int ref_divergence_rhs(const mif_grid *g, const double *u, const double *v, const double *w, double dt,
                       double *out) {
  const auto c = make_constants(g);
  VelocityTensor velocity(*c);
  load_into(velocity.u, u);
  load_into(velocity.v, v);
  load_into(velocity.w, w);
  StaggeredTensor rhs(*c, StaggeringDirection::none);
  load_into(rhs, out);
  STAGGERED_TENSOR_ITERATE_OVER_ALL_OWNER_POINTS(rhs, rhs(i, j, k) = calculate_velocity_divergence(velocity, i, j, k) / dt;)
  store_from(rhs, out);
  return 0;
}

/* TimeVectorFunction::get_difference_over_time(time_1, time_2)
 * (src/VectorFunction.cpp:53-61) of fn, evaluated at n points xyz[3n];
 * out[3n] holds the three components per point. */
int ref_time_difference(ref_eval_fn fn, const void *ctx, double time_1, double time_2, long n, const double *xyz,
                        double *out) {
  const TimeVectorFunction tf = time_function(fn, ctx);
  const VectorFunction d = tf.get_difference_over_time(time_1, time_2);
  for (long q = 0; q < n; q++)
    for (int c = 0; c < 3; c++) out[3 * q + c] = (*d.components[c])(xyz[3 * q], xyz[3 * q + 1], xyz[3 * q + 2]);
  return 0;
}

/* The nhn boundary functions timestep_nhn hands to the pressure solve in each
 * stage: PRESSURE_EQUATION_true(velocity, dt, new_time, prev_time) calls
 * exact_pressure_gradient.get_difference_over_time(new_time, prev_time)
 * (src/Timestep.cpp:89-95) with (time_1, t_n), (time_2, time_1) and
 * (final_time, time_2) in stages 1-3 (:110, :122, :134), where
 * time_1 = t_n + 64/120 dt, time_2 = time_1 + 16/120 dt, final_time =
 * t_n + dt (:99-104). */
int ref_nhn_stage_difference(ref_eval_fn fn, const void *ctx, double dt, double t_n, int stage, long n,
                             const double *xyz, double *out) {
  const Real dt_1 = 64.0 / 120.0 * dt;
  const Real time_1 = t_n + dt_1;
  const Real dt_2 = 16.0 / 120.0 * dt;
  const Real time_2 = time_1 + dt_2;
  const Real final_time = t_n + dt;
  const Real nt = stage == 1 ? time_1 : (stage == 2 ? time_2 : final_time);
  const Real pt = stage == 1 ? t_n : (stage == 2 ? time_1 : time_2);
  return ref_time_difference(fn, ctx, nt, pt, n, xyz, out);
}

/* exact_{u,v,w}_t{1,2} (include/TestCaseBoundaries.h:18-56): test case
 * tcase (1 or 2), component comp, at n points txyz[4n] = (t, x, y, z). */
int ref_lid(int tcase, int comp, long n, const double *txyz, double *out) {
  for (long q = 0; q < n; q++) {
    const double t = txyz[4 * q], x = txyz[4 * q + 1], y = txyz[4 * q + 2], z = txyz[4 * q + 3];
    double r;
    if (tcase == 1)
      r = comp == 0 ? exact_u_t1(t, x, y, z) : (comp == 1 ? exact_v_t1(t, x, y, z) : exact_w_t1(t, x, y, z));
    else
      r = comp == 0 ? exact_u_t2(t, x, y, z) : (comp == 1 ? exact_v_t2(t, x, y, z) : exact_w_t2(t, x, y, z));
    out[q] = r;
  }
  return 0;
}

/* ErrorL1Norm / ErrorL2Norm / ErrorLInfNorm of a velocity (src/Norms.cpp:11-120). */
int ref_norms_velocity(const mif_grid *g, const double *u, const double *v, const double *w, ref_eval_fn fn,
                       const void *ctx, double t, double out[3]) {
  ensure_mpi();
  const auto c = make_constants(g);
  VelocityTensor vel(*c);
  load_into(vel.u, u);
  load_into(vel.v, v);
  load_into(vel.w, w);
  const TimeVectorFunction ev = time_function(fn, ctx);
  out[0] = ErrorL1Norm(vel, ev, t);
  out[1] = ErrorL2Norm(vel, ev, t);
  out[2] = ErrorLInfNorm(vel, ev, t);
  return 0;
}

/* The same norms of a pressure field against component 0 of fn
 * (src/Norms.cpp:122-163). */
int ref_norms_pressure(const mif_grid *g, const double *p, ref_eval_fn fn, const void *ctx, double t,
                       double out[3]) {
  ensure_mpi();
  const auto c = make_constants(g);
  StaggeredTensor pt(*c, StaggeringDirection::none);
  load_into(pt, p);
  const std::function<Real(Real, Real, Real, Real)> fp = [fn, ctx](Real tt, Real x, Real y, Real z) {
    return fn(ctx, 0, tt, x, y, z);
  };
  out[0] = ErrorL1Norm(pt, fp, t);
  out[1] = ErrorL2Norm(pt, fp, t);
  out[2] = ErrorLInfNorm(pt, fp, t);
  return 0;
}

/* Tensor::dump (include/Tensor.h:465-476) of a field of staggering stag. */
int ref_dump(const mif_grid *g, int stag, const double *f, const char *path) {
  const auto c = make_constants(g);
  if (stag == 3) {
    StaggeredTensor p(*c, StaggeringDirection::none);
    load_into(p, f);
    return p.dump(path) ? 0 : 1;
  }
  VelocityTensor vel(*c);
  StaggeredTensor &t3 = component(vel, stag);
  load_into(t3, f);
  return t3.dump(path) ? 0 : 1;
}

}  // extern "C"
