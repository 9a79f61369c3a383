/*
 * mif_oracle.c -- CPU oracle (test infrastructure, see mif_oracle.h).
 *
 * A single-rank restatement of the reference solver's hot path.  Every
 * function cites the reference file:line it follows.  Loop orders and the
 * grouping of floating-point operations follow the reference so that the
 * oracle agrees with it to rounding; the GPU library is checked against this
 * file at 1e-11 relative.  Built without fast-math and without FMA
 * contraction (oracle/Makefile).
 */
#include "mif_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Grid constants: src/Constants.cpp:58-120 with Py = Pz = 1, rank 0.        */
/* ------------------------------------------------------------------------ */
int orc_grid_init(orc_grid *g, const mif_grid *d) {
  memset(g, 0, sizeof(*g));
  const uint64_t Ng[3] = {d->Nx_global, d->Ny_global, d->Nz_global};
  const double L[3] = {d->x_size, d->y_size, d->z_size};
  const double mn[3] = {d->min_x, d->min_y, d->min_z};
  if (d->num_time_steps == 0 || !(d->final_time > 0)) return MIF_ERR_INVALID_ARGUMENT;
  g->Re = d->Re;
  g->dt = d->final_time / d->num_time_steps; /* src/Constants.cpp:69 */
  for (int a = 0; a < 3; a++) {
    if (Ng[a] < 2) return MIF_ERR_INVALID_ARGUMENT;
    g->Ng[a] = (long)Ng[a];
    g->periodic[a] = d->periodic[a] ? 1 : 0;
    g->L[a] = L[a];
    g->min[a] = mn[a];
    /* Nx/Ny/Nz at P = 1 (src/Constants.cpp:80-82) */
    g->N[a] = g->Ng[a] + g->periodic[a];
    /* base_i/j/k (src/Constants.cpp:92-94) */
    g->base[a] = g->periodic[a] ? -1 : 0;
    /* dx = size / (N_global - 1) (src/Constants.cpp:70) */
    const double h = L[a] / (double)(g->Ng[a] - 1);
    g->h[a] = h;
    g->one_over_2h[a] = 1 / (2 * h);
    g->one_over_8h[a] = 1 / (8 * h);
    g->one_over_h2Re[a] = 1 / (d->Re * h * h);
    g->h_over_2[a] = h / 2;
    g->one_over_h[a] = 1 / h;
    /* Nx_points... (src/PressureSolverStructures.cpp:15-17) */
    g->npts[a] = g->Ng[a] - (g->periodic[a] ? 1 : 0);
    if (g->N[a] < 2) return MIF_ERR_INVALID_ARGUMENT;
  }
  return MIF_OK;
}

/* Extents per staggering (src/StaggeredTensor.cpp:5-9; Nx_staggered =
 * periodic ? Nx : Nx + 1, Ny/Nz_staggered = Ny/Nz + 1 on the last rank,
 * src/Constants.cpp:83-85). */
void orc_extents(const orc_grid *g, int s, long e[3]) {
  e[0] = g->N[0];
  e[1] = g->N[1];
  e[2] = g->N[2];
  if (s == 0) e[0] = g->periodic[0] ? g->N[0] : g->N[0] + 1;
  if (s == 1) e[1] = g->N[1] + 1;
  if (s == 2) e[2] = g->N[2] + 1;
}

long orc_field_size(const orc_grid *g, int s) {
  long e[3];
  orc_extents(g, s, e);
  return e[0] * e[1] * e[2];
}

#define AT(f, e, i, j, k) ((f)[(i) + (e)[0] * ((j) + (e)[1] * (k))])

/* ------------------------------------------------------------------------ */
/* Analytic functions.                                                       */
/* ------------------------------------------------------------------------ */

/* Ethier-Steinman u (generators/manufsol.py:36-44); v and w are u with the
 * arguments permuted cyclically (x,y,z) -> (y,z,x) -> (z,x,y). */
static double es_u(double t, double x, double y, double z, double Re) {
  const double a = M_PI / 4, d = M_PI / 2;
  return -a * (exp(a * x) * sin(a * y + d * z) + exp(a * z) * cos(a * x + d * y)) *
         exp(-d * d * t / Re);
}

/* Ethier-Steinman p (generators/manufsol.py:55-68). */
static double es_p(double t, double x, double y, double z, double Re) {
  const double a = M_PI / 4, d = M_PI / 2;
  const double s =
      exp(2 * a * x) + exp(2 * a * y) + exp(2 * a * z) +
      2 * sin(a * x + d * y) * cos(a * z + d * x) * exp(a * (y + z)) +
      2 * sin(a * y + d * z) * cos(a * x + d * y) * exp(a * (z + x)) +
      2 * sin(a * z + d * x) * cos(a * y + d * z) * exp(a * (x + y));
  return -a * a / 2 * s * exp(-2 * d * d * t / Re);
}

/* d p / d x of the Ethier-Steinman pressure, differentiated term by term;
 * d p/d y and d p/d z follow from the cyclic symmetry of p. */
static double es_dpdx(double t, double x, double y, double z, double Re) {
  const double a = M_PI / 4, d = M_PI / 2;
  const double S1 = sin(a * x + d * y), C1 = cos(a * z + d * x);
  const double S2 = sin(a * y + d * z), C2 = cos(a * x + d * y);
  const double S3 = sin(a * z + d * x), C3 = cos(a * y + d * z);
  const double s = 2 * a * exp(2 * a * x) +
                   2 * exp(a * (y + z)) * (a * C2 * C1 - d * S1 * S3) +
                   2 * a * exp(a * (z + x)) * S2 * (C2 - S1) +
                   2 * exp(a * (x + y)) * C3 * (d * C1 + a * S3);
  return -a * a / 2 * s * exp(-2 * d * d * t / Re);
}

/* Velocity-only field (generators/manufsol_velocity.py:55-57). */
static double vt_vel(int c, double t, double x, double y, double z) {
  if (c == 0) return sin(x) * cos(y) * sin(z) * sin(t);
  if (c == 1) return cos(x) * sin(y) * sin(z) * sin(t);
  return 2 * cos(x) * cos(y) * cos(z) * sin(t);
}

/* Its forcing f = u_t + (u.grad)u - lap(u)/Re (generators/
 * manufsol_velocity.py:40-50), simplified by hand (derivation in
 * oracle/derive_manufactured.py). */
static double vt_forcing(int c, double t, double x, double y, double z, double Re) {
  const double st = sin(t), ct = cos(t);
  const double sx = sin(x), cx = cos(x), sy = sin(y), cy = cos(y), sz = sin(z), cz = cos(z);
  if (c == 0)
    return sx * cx * st * st * (2 - 2 * sy * sy - sz * sz) + sx * sz * cy * (ct + 3 * st / Re);
  if (c == 1)
    return sy * cy * st * st * (2 - 2 * sx * sx - sz * sz) + sy * sz * cx * (ct + 3 * st / Re);
  return sz * cz * st * st * (2 * sx * sx + 2 * sy * sy - 4) + cx * cy * cz * (2 * ct + 6 * st / Re);
}

/* Poisson test field (generators/manufsol_pressure.py:36-39). */
static double pt_vel(int c, double t, double x, double y, double z) {
  if (c == 0) return -sin(x) * cos(y) * cos(z) * t;
  if (c == 1) return -cos(x) * sin(y) * cos(z) * t;
  return -cos(x) * cos(y) * sin(z) * t;
}

/* Lid-driven boundary data (include/TestCaseBoundaries.h:18-56). */
static double lid(int c, double x, double xlid) {
  const double eps = 1e-12;
  if (c != 1) return 0.0;
  if (x < xlid + eps && x > xlid - eps) return 1.0;
  return 0.0;
}

static double fn_base(int kind, int c, double t, double x, double y, double z, const mif_function *f) {
  const double Re = f->Re;
  switch (kind) {
    case MIF_FN_ZERO: return 0.0;
    case MIF_FN_ES_VELOCITY:
      if (c == 0) return es_u(t, x, y, z, Re);
      if (c == 1) return es_u(t, y, z, x, Re);
      return es_u(t, z, x, y, Re);
    case MIF_FN_ES_PRESSURE: return es_p(t, x, y, z, Re);
    case MIF_FN_ES_GRAD_P:
      if (c == 0) return es_dpdx(t, x, y, z, Re);
      if (c == 1) return es_dpdx(t, y, z, x, Re);
      return es_dpdx(t, z, x, y, Re);
    case MIF_FN_VT_VELOCITY: return vt_vel(c, t, x, y, z);
    case MIF_FN_VT_FORCING: return vt_forcing(c, t, x, y, z, Re);
    case MIF_FN_PT_VELOCITY: return pt_vel(c, t, x, y, z);
    case MIF_FN_PT_PRESSURE: return cos(x) * cos(y) * cos(z) * t;
    case MIF_FN_PT_GRAD_P: return pt_vel(c, t, x, y, z);
    case MIF_FN_LID_1: return lid(c, x, 1.0);
    case MIF_FN_LID_2: return lid(c, x, -0.5);
    case MIF_FN_CONSTANT: return f->c[c];
    default: return 0.0;
  }
}

double orc_fn_eval(const mif_function *f, int c, double t, double x, double y, double z) {
  if (f->kind == MIF_FN_TIME_DIFFERENCE)
    return fn_base(f->base, c, f->t1, x, y, z, f) - fn_base(f->base, c, f->t0, x, y, z, f);
  return fn_base(f->kind, c, t, x, y, z, f);
}

/* Coordinates of index (i,j,k) of a field with staggering s
 * (include/StaggeredTensor.h:104-110, include/VelocityTensor.h:16-78):
 * min + h*(base+i), minus h/2 along the staggered axis. */
static inline void coords(const orc_grid *g, int s, long i, long j, long k, double *x) {
  const long idx[3] = {i, j, k};
  for (int a = 0; a < 3; a++) {
    x[a] = g->min[a] + g->h[a] * (double)(g->base[a] + idx[a]);
    if (a == s) x[a] = x[a] - g->h_over_2[a];
  }
}

static inline double eval_at(const orc_grid *g, int s, const mif_function *f, int c, double t,
                             long i, long j, long k) {
  double x[3];
  coords(g, s, i, j, k, x);
  return orc_fn_eval(f, c, t, x[0], x[1], x[2]);
}

/* StaggeredTensor::set (src/StaggeredTensor.cpp:205-219). */
void orc_set(const orc_grid *g, double *f, int s, const mif_function *fn, int comp, double t,
             int include_border) {
  long e[3];
  orc_extents(g, s, e);
  const long lo = include_border ? 0 : 1;
  const long hi[3] = {include_border ? e[0] : e[0] - 1, include_border ? e[1] : e[1] - 1,
                      include_border ? e[2] : e[2] - 1};
#pragma omp parallel for schedule(static)
  for (long k = lo; k < hi[2]; k++)
    for (long j = lo; j < hi[1]; j++)
      for (long i = lo; i < hi[0]; i++) AT(f, e, i, j, k) = eval_at(g, s, fn, comp, t, i, j, k);
}

/* StaggeredTensor::apply_periodic_bc (src/StaggeredTensor.cpp:221-256). */
void orc_apply_periodic(const orc_grid *g, double *f, int s) {
  long e[3];
  orc_extents(g, s, e);
  if (g->periodic[0])
    for (long k = 0; k < e[2]; k++)
      for (long j = 0; j < e[1]; j++) {
        AT(f, e, 0, j, k) = AT(f, e, e[0] - 2, j, k);
        AT(f, e, e[0] - 1, j, k) = AT(f, e, 1, j, k);
      }
  if (g->periodic[1])
    for (long k = 0; k < e[2]; k++)
      for (long i = 0; i < e[0]; i++) {
        AT(f, e, i, 0, k) = AT(f, e, i, e[1] - 2, k);
        AT(f, e, i, e[1] - 1, k) = AT(f, e, i, 1, k);
      }
  if (g->periodic[2])
    for (long j = 0; j < e[1]; j++)
      for (long i = 0; i < e[0]; i++) {
        AT(f, e, i, j, 0) = AT(f, e, i, j, e[2] - 2);
        AT(f, e, i, j, e[2] - 1) = AT(f, e, i, j, 1);
      }
}

/* Unstaggered evaluation (evaluate_function_at_index_unstaggered,
 * include/StaggeredTensor.h:95-101). */
static inline double eval_node(const orc_grid *g, const mif_function *f, int c, double t, long i,
                               long j, long k) {
  return eval_at(g, 3, f, c, t, i, j, k);
}

/* VelocityTensor::apply_bc (src/VelocityTensor.cpp:36-233): faces z-, z+,
 * y-, y+, x-, x+ on non-periodic axes; the normal component at a face is the
 * exact value at the node plus/minus h/2 times the tangential divergence of
 * the exact field; then the periodic ghost copies. */
void orc_apply_bc(const orc_grid *g, double *u, double *v, double *w, const mif_function *fn,
                  double t) {
  double *T[3] = {u, v, w};
  const long Nx = g->N[0], Ny = g->N[1], Nz = g->N[2];
  for (int c = 0; c < 3; c++) {
    double *f = T[c];
    long e[3];
    orc_extents(g, c, e);
    /* Face 1: z = z_min (src/VelocityTensor.cpp:48-73) */
    if (!g->periodic[2]) {
      for (long j = 0; j < e[1]; j++)
        for (long i = 0; i < e[0]; i++) {
          if (c == 2) {
            const double wb = eval_node(g, fn, 2, t, i, j, 0);
            const double du_dx =
                (eval_at(g, 0, fn, 0, t, i + 1, j, 0) - eval_at(g, 0, fn, 0, t, i, j, 0)) * g->one_over_h[0];
            const double dv_dy =
                (eval_at(g, 1, fn, 1, t, i, j + 1, 0) - eval_at(g, 1, fn, 1, t, i, j, 0)) * g->one_over_h[1];
            AT(f, e, i, j, 0) = wb + g->h_over_2[2] * (du_dx + dv_dy);
          } else {
            AT(f, e, i, j, 0) = eval_at(g, c, fn, c, t, i, j, 0);
          }
        }
      /* Face 2: z = z_max (src/VelocityTensor.cpp:76-106) */
      for (long j = 0; j < e[1]; j++)
        for (long i = 0; i < e[0]; i++) {
          if (c == 2) {
            const double wb = eval_node(g, fn, 2, t, i, j, Nz - 1);
            const double du_dx = (eval_at(g, 0, fn, 0, t, i + 1, j, Nz - 1) -
                                  eval_at(g, 0, fn, 0, t, i, j, Nz - 1)) * g->one_over_h[0];
            const double dv_dy = (eval_at(g, 1, fn, 1, t, i, j + 1, Nz - 1) -
                                  eval_at(g, 1, fn, 1, t, i, j, Nz - 1)) * g->one_over_h[1];
            AT(f, e, i, j, e[2] - 1) = wb - g->h_over_2[2] * (du_dx + dv_dy);
          } else {
            AT(f, e, i, j, Nz - 1) = eval_at(g, c, fn, c, t, i, j, Nz - 1);
          }
        }
    }
    if (!g->periodic[1]) {
      /* Face 3: y = y_min (src/VelocityTensor.cpp:109-131) */
      for (long k = 0; k < e[2]; k++)
        for (long i = 0; i < e[0]; i++) {
          if (c == 1) {
            const double vb = eval_node(g, fn, 1, t, i, 0, k);
            const double du_dx =
                (eval_at(g, 0, fn, 0, t, i + 1, 0, k) - eval_at(g, 0, fn, 0, t, i, 0, k)) * g->one_over_h[0];
            const double dw_dz =
                (eval_at(g, 2, fn, 2, t, i, 0, k + 1) - eval_at(g, 2, fn, 2, t, i, 0, k)) * g->one_over_h[2];
            AT(f, e, i, 0, k) = vb + g->h_over_2[1] * (du_dx + dw_dz);
          } else {
            AT(f, e, i, 0, k) = eval_at(g, c, fn, c, t, i, 0, k);
          }
        }
      /* Face 4: y = y_max (src/VelocityTensor.cpp:134-164) */
      for (long k = 0; k < e[2]; k++)
        for (long i = 0; i < e[0]; i++) {
          if (c == 1) {
            const double vb = eval_node(g, fn, 1, t, i, Ny - 1, k);
            const double du_dx = (eval_at(g, 0, fn, 0, t, i + 1, Ny - 1, k) -
                                  eval_at(g, 0, fn, 0, t, i, Ny - 1, k)) * g->one_over_h[0];
            const double dw_dz = (eval_at(g, 2, fn, 2, t, i, Ny - 1, k + 1) -
                                  eval_at(g, 2, fn, 2, t, i, Ny - 1, k)) * g->one_over_h[2];
            AT(f, e, i, e[1] - 1, k) = vb - g->h_over_2[1] * (du_dx + dw_dz);
          } else {
            AT(f, e, i, Ny - 1, k) = eval_at(g, c, fn, c, t, i, Ny - 1, k);
          }
        }
    }
    if (!g->periodic[0]) {
      /* Face 5: x = x_min (src/VelocityTensor.cpp:167-189) */
      for (long k = 0; k < e[2]; k++)
        for (long j = 0; j < e[1]; j++) {
          if (c == 0) {
            const double ub = eval_node(g, fn, 0, t, 0, j, k);
            const double dv_dy =
                (eval_at(g, 1, fn, 1, t, 0, j + 1, k) - eval_at(g, 1, fn, 1, t, 0, j, k)) * g->one_over_h[1];
            const double dw_dz =
                (eval_at(g, 2, fn, 2, t, 0, j, k + 1) - eval_at(g, 2, fn, 2, t, 0, j, k)) * g->one_over_h[2];
            AT(f, e, 0, j, k) = ub + g->h_over_2[0] * (dv_dy + dw_dz);
          } else {
            AT(f, e, 0, j, k) = eval_at(g, c, fn, c, t, 0, j, k);
          }
        }
      /* Face 6: x = x_max (src/VelocityTensor.cpp:192-222); the node index of
       * the wall is Nx_domains = Nx_global - 1. */
      for (long k = 0; k < e[2]; k++)
        for (long j = 0; j < e[1]; j++) {
          if (c == 0) {
            const double ub = eval_node(g, fn, 0, t, g->Ng[0] - 1, j, k);
            const double dv_dy = (eval_at(g, 1, fn, 1, t, Nx - 1, j + 1, k) -
                                  eval_at(g, 1, fn, 1, t, Nx - 1, j, k)) * g->one_over_h[1];
            const double dw_dz = (eval_at(g, 2, fn, 2, t, Nx - 1, j, k + 1) -
                                  eval_at(g, 2, fn, 2, t, Nx - 1, j, k)) * g->one_over_h[2];
            AT(f, e, e[0] - 1, j, k) = ub - g->h_over_2[0] * (dv_dy + dw_dz);
          } else {
            AT(f, e, Nx - 1, j, k) = eval_at(g, c, fn, c, t, Nx - 1, j, k);
          }
        }
    }
    /* periodic ghost copies (src/VelocityTensor.cpp:225) */
    orc_apply_periodic(g, f, c);
  }
}

/* ------------------------------------------------------------------------ */
/* Stencils: include/MomentumEquation.h:43-255 (OPT_CPU_CACHE branch, the    */
/* one the reference builds, Makefile:29), include/PressureGradient.h:9-25,  */
/* include/VelocityDivergence.h:9-20.                                        */
/* ------------------------------------------------------------------------ */
double orc_momentum_rhs(const orc_grid *g, double *const vel[3], int comp, long i, long j, long k) {
  long eu[3], ev[3], ew[3];
  orc_extents(g, 0, eu);
  orc_extents(g, 1, ev);
  orc_extents(g, 2, ew);
  const double *u = vel[0], *v = vel[1], *w = vel[2];
  if (comp == 0) {
    const double u13 = AT(u, eu, i - 1, j, k), u11 = AT(u, eu, i, j - 1, k), u5 = AT(u, eu, i, j, k - 1),
                 u14 = AT(u, eu, i, j, k), u23 = AT(u, eu, i, j, k + 1), u17 = AT(u, eu, i, j + 1, k),
                 u15 = AT(u, eu, i + 1, j, k);
    const double v13 = AT(v, ev, i - 1, j, k), v16 = AT(v, ev, i - 1, j + 1, k), v14 = AT(v, ev, i, j, k),
                 v17 = AT(v, ev, i, j + 1, k);
    const double w13 = AT(w, ew, i - 1, j, k), w22 = AT(w, ew, i - 1, j, k + 1), w14 = AT(w, ew, i, j, k),
                 w23 = AT(w, ew, i, j, k + 1);
    const double conv = -u14 * (u15 - u13) * g->one_over_2h[0] -
                        (v17 + v14 + v16 + v13) * (u17 - u11) * g->one_over_8h[1] -
                        (w23 + w14 + w22 + w13) * (u23 - u5) * g->one_over_8h[2];
    const double diff = (u15 - 2 * u14 + u13) * g->one_over_h2Re[0] +
                        (u17 - 2 * u14 + u11) * g->one_over_h2Re[1] +
                        (u23 - 2 * u14 + u5) * g->one_over_h2Re[2];
    return conv + diff;
  }
  if (comp == 1) {
    const double u11 = AT(u, eu, i, j - 1, k), u14 = AT(u, eu, i, j, k), u12 = AT(u, eu, i + 1, j - 1, k),
                 u15 = AT(u, eu, i + 1, j, k);
    const double v13 = AT(v, ev, i - 1, j, k), v11 = AT(v, ev, i, j - 1, k), v5 = AT(v, ev, i, j, k - 1),
                 v14 = AT(v, ev, i, j, k), v23 = AT(v, ev, i, j, k + 1), v17 = AT(v, ev, i, j + 1, k),
                 v15 = AT(v, ev, i + 1, j, k);
    const double w11 = AT(w, ew, i, j - 1, k), w20 = AT(w, ew, i, j - 1, k + 1), w14 = AT(w, ew, i, j, k),
                 w23 = AT(w, ew, i, j, k + 1);
    const double conv = -(u15 + u14 + u12 + u11) * (v15 - v13) * g->one_over_8h[0] -
                        v14 * (v17 - v11) * g->one_over_2h[1] -
                        (w23 + w14 + w20 + w11) * (v23 - v5) * g->one_over_8h[2];
    const double diff = (v15 - 2 * v14 + v13) * g->one_over_h2Re[0] +
                        (v17 - 2 * v14 + v11) * g->one_over_h2Re[1] +
                        (v23 - 2 * v14 + v5) * g->one_over_h2Re[2];
    return conv + diff;
  }
  {
    const double u5 = AT(u, eu, i, j, k - 1), u14 = AT(u, eu, i, j, k), u6 = AT(u, eu, i + 1, j, k - 1),
                 u15 = AT(u, eu, i + 1, j, k);
    const double v5 = AT(v, ev, i, j, k - 1), v14 = AT(v, ev, i, j, k), v8 = AT(v, ev, i, j + 1, k - 1),
                 v17 = AT(v, ev, i, j + 1, k);
    const double w13 = AT(w, ew, i - 1, j, k), w11 = AT(w, ew, i, j - 1, k), w5 = AT(w, ew, i, j, k - 1),
                 w14 = AT(w, ew, i, j, k), w23 = AT(w, ew, i, j, k + 1), w17 = AT(w, ew, i, j + 1, k),
                 w15 = AT(w, ew, i + 1, j, k);
    const double conv = -(u15 + u14 + u6 + u5) * (w15 - w13) * g->one_over_8h[0] -
                        (v17 + v14 + v8 + v5) * (w17 - w11) * g->one_over_8h[1] -
                        w14 * (w23 - w5) * g->one_over_2h[2];
    const double diff = (w15 - 2 * w14 + w13) * g->one_over_h2Re[0] +
                        (w17 - 2 * w14 + w11) * g->one_over_h2Re[1] +
                        (w23 - 2 * w14 + w5) * g->one_over_h2Re[2];
    return conv + diff;
  }
}

/* pressure_gradient_{u,v,w} (include/PressureGradient.h:9-25). */
static inline double pgrad(const orc_grid *g, const double *p, int comp, long i, long j, long k) {
  long e[3];
  orc_extents(g, 3, e);
  if (comp == 0) return (AT(p, e, i, j, k) - AT(p, e, i - 1, j, k)) * g->one_over_h[0];
  if (comp == 1) return (AT(p, e, i, j, k) - AT(p, e, i, j - 1, k)) * g->one_over_h[1];
  return (AT(p, e, i, j, k) - AT(p, e, i, j, k - 1)) * g->one_over_h[2];
}

/* calculate_velocity_divergence (include/VelocityDivergence.h:9-20). */
static inline double divergence(const orc_grid *g, double *const vel[3], long i, long j, long k) {
  long eu[3], ev[3], ew[3];
  orc_extents(g, 0, eu);
  orc_extents(g, 1, ev);
  orc_extents(g, 2, ew);
  const double du_dx = (AT(vel[0], eu, i + 1, j, k) - AT(vel[0], eu, i, j, k)) * g->one_over_h[0];
  const double dv_dy = (AT(vel[1], ev, i, j + 1, k) - AT(vel[1], ev, i, j, k)) * g->one_over_h[1];
  const double dw_dz = (AT(vel[2], ew, i, j, k + 1) - AT(vel[2], ew, i, j, k)) * g->one_over_h[2];
  return du_dx + dv_dy + dw_dz;
}

/* ------------------------------------------------------------------------ */
/* Transforms.  Complex mixed-radix FFT (decimation in time), used to build  */
/* FFTW's R2HC / HC2R / REDFT00 from their documented definitions.           */
/* ------------------------------------------------------------------------ */
typedef struct { double re, im; } cplx;

typedef struct {
  long n;
  cplx *w; /* w[e] = exp(-2 pi i e / n) */
} fft_plan;

static void plan_init(fft_plan *P, long n) {
  P->n = n;
  P->w = (cplx *)malloc(sizeof(cplx) * (size_t)(n > 0 ? n : 1));
  for (long e = 0; e < n; e++) {
    const double ang = 2.0 * M_PI * (double)e / (double)n;
    P->w[e].re = cos(ang);
    P->w[e].im = -sin(ang);
  }
}
static void plan_free(fft_plan *P) { free(P->w); }

static long smallest_factor(long n) {
  if (n % 2 == 0) return 2;
  for (long p = 3; p * p <= n; p += 2)
    if (n % p == 0) return p;
  return n;
}

/* out[k] = sum_j in[j*stride] * exp(-2 pi i j k / n), k < n; tw = P->n / n. */
static void fft_rec(const fft_plan *P, const cplx *in, long stride, cplx *out, long n, long tw) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  const long p = smallest_factor(n), m = n / p;
  for (long r = 0; r < p; r++) fft_rec(P, in + r * stride, stride * p, out + r * m, m, tw * p);
  cplx tbuf[64];
  cplx *t = p <= 64 ? tbuf : (cplx *)malloc(sizeof(cplx) * (size_t)p);
  cplx obuf[64];
  cplx *o = p <= 64 ? obuf : (cplx *)malloc(sizeof(cplx) * (size_t)p);
  for (long k = 0; k < m; k++) {
    for (long r = 0; r < p; r++) {
      const cplx f = out[r * m + k];
      const cplx wr = P->w[((r * k) % n) * tw];
      t[r].re = f.re * wr.re - f.im * wr.im;
      t[r].im = f.re * wr.im + f.im * wr.re;
    }
    for (long s = 0; s < p; s++) {
      double re = 0, im = 0;
      for (long r = 0; r < p; r++) {
        const cplx wp = P->w[((r * s) % p) * m * tw];
        re += t[r].re * wp.re - t[r].im * wp.im;
        im += t[r].re * wp.im + t[r].im * wp.re;
      }
      o[s].re = re;
      o[s].im = im;
    }
    for (long s = 0; s < p; s++) out[k + m * s] = o[s];
  }
  if (p > 64) {
    free(t);
    free(o);
  }
}

/* Forward DFT of n complex values (sign -1), out of place. */
static void cfft(const fft_plan *P, const cplx *in, cplx *out) { fft_rec(P, in, 1, out, P->n, 1); }

/* FFTW_R2HC: r_0..r_{n/2}, i_{(n+1)/2-1}..i_1 (halfcomplex order). */
static void r2hc_plan(const fft_plan *P, const double *in, double *out, cplx *a, cplx *b) {
  const long n = P->n;
  for (long j = 0; j < n; j++) {
    a[j].re = in[j];
    a[j].im = 0;
  }
  cfft(P, a, b);
  for (long k = 0; k <= n / 2; k++) out[k] = b[k].re;
  for (long k = 1; k < (n + 1) / 2; k++) out[n - k] = b[k].im;
}

/* FFTW_HC2R: x_j = sum_k X_k exp(+2 pi i j k / n), X Hermitian, unnormalised.
 * Computed as conj(FFT(conj X)). */
static void hc2r_plan(const fft_plan *P, const double *in, double *out, cplx *a, cplx *b) {
  const long n = P->n;
  a[0].re = in[0];
  a[0].im = 0;
  for (long k = 1; k < (n + 1) / 2; k++) {
    a[k].re = in[k];
    a[k].im = -in[n - k]; /* conj */
    a[n - k].re = in[k];
    a[n - k].im = in[n - k];
  }
  if (n % 2 == 0) {
    a[n / 2].re = in[n / 2];
    a[n / 2].im = 0;
  }
  cfft(P, a, b);
  for (long j = 0; j < n; j++) out[j] = b[j].re;
}

/* FFTW_REDFT00 on n points: Y_k = X_0 + (-1)^k X_{n-1} + 2 sum_{j=1}^{n-2}
 * X_j cos(pi j k / (n-1)), via the DFT of the even extension of length
 * 2(n-1) (P->n == 2(n-1)). */
static void redft00_plan(const fft_plan *P, long n, const double *in, double *out, cplx *a, cplx *b) {
  const long M = P->n;
  for (long j = 0; j < n; j++) {
    a[j].re = in[j];
    a[j].im = 0;
  }
  for (long j = 1; j < n - 1; j++) {
    a[M - j].re = in[j];
    a[M - j].im = 0;
  }
  cfft(P, a, b);
  for (long k = 0; k < n; k++) out[k] = b[k].re;
}

void orc_r2hc(long n, const double *in, double *out) {
  fft_plan P;
  plan_init(&P, n);
  cplx *a = malloc(sizeof(cplx) * n), *b = malloc(sizeof(cplx) * n);
  r2hc_plan(&P, in, out, a, b);
  free(a);
  free(b);
  plan_free(&P);
}
void orc_hc2r(long n, const double *in, double *out) {
  fft_plan P;
  plan_init(&P, n);
  cplx *a = malloc(sizeof(cplx) * n), *b = malloc(sizeof(cplx) * n);
  hc2r_plan(&P, in, out, a, b);
  free(a);
  free(b);
  plan_free(&P);
}
void orc_redft00(long n, const double *in, double *out) {
  if (n < 2) {
    out[0] = in[0];
    return;
  }
  fft_plan P;
  plan_init(&P, 2 * (n - 1));
  cplx *a = malloc(sizeof(cplx) * P.n), *b = malloc(sizeof(cplx) * P.n);
  redft00_plan(&P, n, in, out, a, b);
  free(a);
  free(b);
  plan_free(&P);
}

void orc_r2hc_direct(long n, const double *in, double *out) {
  for (long k = 0; k <= n / 2; k++) {
    double re = 0;
    for (long j = 0; j < n; j++) re += in[j] * cos(2.0 * M_PI * (double)((j * k) % n) / (double)n);
    out[k] = re;
  }
  for (long k = 1; k < (n + 1) / 2; k++) {
    double im = 0;
    for (long j = 0; j < n; j++) im -= in[j] * sin(2.0 * M_PI * (double)((j * k) % n) / (double)n);
    out[n - k] = im;
  }
}
void orc_hc2r_direct(long n, const double *in, double *out) {
  for (long j = 0; j < n; j++) {
    double s = in[0];
    for (long k = 1; k < (n + 1) / 2; k++) {
      const double ang = 2.0 * M_PI * (double)((j * k) % n) / (double)n;
      s += 2 * (in[k] * cos(ang) - in[n - k] * sin(ang));
    }
    if (n % 2 == 0) s += in[n / 2] * ((j % 2) ? -1.0 : 1.0);
    out[j] = s;
  }
}
void orc_redft00_direct(long n, const double *in, double *out) {
  for (long k = 0; k < n; k++) {
    double s = in[0] + ((k % 2) ? -1.0 : 1.0) * in[n - 1];
    for (long j = 1; j < n - 1; j++)
      s += 2 * in[j] * cos(M_PI * (double)((j * k) % (2 * (n - 1))) / (double)(n - 1));
    out[k] = s;
  }
}

/* ------------------------------------------------------------------------ */
/* Spectral Poisson solve (src/PressureEquation.cpp:65-264,                   */
/* src/PressureSolverStructures.cpp:5-69).  The reference transforms along x, */
/* transposes so y is contiguous, transforms, transposes so z is contiguous;  */
/* at one rank the transposes are permutations, so each 1-D transform is      */
/* applied to the same line of data whichever layout holds it.  The oracle   */
/* keeps the x-fastest layout and gathers strided lines.                      */
/* ------------------------------------------------------------------------ */

/* compute_eigenvalue_neumann / _periodic (src/PressureSolverStructures.cpp:5-11) */
static double eig_neumann(long index, double delta, long N) {
  return 2.0 * (cos(M_PI * (double)index / (double)(N - 1)) - 1.0) / (delta * delta);
}
static double eig_periodic(long index, double delta, long N) {
  return 2.0 * (cos(2.0 * M_PI * (double)index / (double)N) - 1.0) / (delta * delta);
}

typedef enum { AX_FWD, AX_INV } ax_dir;

/* Per-line transforms by O(N^2) direct sums instead of the FFT: a second
 * FP64 evaluation order of the same solve (orc_poisson_packed_direct). */
static int g_direct_transforms = 0;

/* Transform every line of b (extents n[3], x fastest) along axis a. */
static void transform_axis(const orc_grid *g, double *b, const long n[3], int a, ax_dir dir,
                           double norm) {
  const long len = n[a];
  const long stride = a == 0 ? 1 : (a == 1 ? n[0] : n[0] * n[1]);
  const int per = g->periodic[a];
  const int b1 = a == 0 ? 1 : 0, b2 = a == 2 ? 1 : 2;
  const long nl1 = n[b1], nl2 = n[b2];
  const long s1 = b1 == 0 ? 1 : n[0], s2 = b2 == 1 ? n[0] : n[0] * n[1];
  fft_plan P;
  if (per)
    plan_init(&P, len);
  else
    plan_init(&P, len >= 2 ? 2 * (len - 1) : 1);
#pragma omp parallel
  {
    double *lin = malloc(sizeof(double) * len), *lout = malloc(sizeof(double) * len);
    cplx *ca = malloc(sizeof(cplx) * P.n), *cb = malloc(sizeof(cplx) * P.n);
#pragma omp for collapse(2) schedule(static)
    for (long q2 = 0; q2 < nl2; q2++)
      for (long q1 = 0; q1 < nl1; q1++) {
        double *base = b + q1 * s1 + q2 * s2;
        for (long t = 0; t < len; t++) lin[t] = base[t * stride];
        if (g_direct_transforms) {
          if (per)
            (dir == AX_FWD ? orc_r2hc_direct : orc_hc2r_direct)(len, lin, lout);
          else if (len < 2)
            lout[0] = lin[0];
          else
            orc_redft00_direct(len, lin, lout);
        } else if (per) {
          if (dir == AX_FWD)
            r2hc_plan(&P, lin, lout, ca, cb);
          else
            hc2r_plan(&P, lin, lout, ca, cb);
        } else {
          if (len < 2)
            lout[0] = lin[0];
          else
            redft00_plan(&P, len, lin, lout, ca, cb);
        }
        if (dir == AX_FWD)
          for (long t = 0; t < len; t++) base[t * stride] = lout[t];
        else /* divide after each inverse axis (src/PressureEquation.cpp:177,210,244) */
          for (long t = 0; t < len; t++) base[t * stride] = lout[t] / norm;
      }
    free(lin);
    free(lout);
    free(ca);
    free(cb);
  }
  plan_free(&P);
}

void orc_poisson_packed(const orc_grid *g, double *b) {
  const long *n = g->npts;
  /* forward x, y, z (src/PressureEquation.cpp:79-154) */
  transform_axis(g, b, n, 0, AX_FWD, 1.0);
  transform_axis(g, b, n, 1, AX_FWD, 1.0);
  transform_axis(g, b, n, 2, AX_FWD, 1.0);
  /* multiply by 1/(lx + ly + lz) and zero the constant mode
   * (src/PressureEquation.cpp:158-163, src/PressureSolverStructures.cpp:49-68) */
  double *lx = malloc(sizeof(double) * n[0]), *ly = malloc(sizeof(double) * n[1]),
         *lz = malloc(sizeof(double) * n[2]);
  for (long i = 0; i < n[0]; i++)
    lx[i] = g->periodic[0] ? eig_periodic(i, g->h[0], n[0]) : eig_neumann(i, g->h[0], n[0]);
  for (long j = 0; j < n[1]; j++)
    ly[j] = g->periodic[1] ? eig_periodic(j, g->h[1], n[1]) : eig_neumann(j, g->h[1], n[1]);
  for (long k = 0; k < n[2]; k++)
    lz[k] = g->periodic[2] ? eig_periodic(k, g->h[2], n[2]) : eig_neumann(k, g->h[2], n[2]);
#pragma omp parallel for schedule(static)
  for (long k = 0; k < n[2]; k++)
    for (long j = 0; j < n[1]; j++)
      for (long i = 0; i < n[0]; i++) {
        const double ev = 1 / (lx[i] + ly[j] + lz[k]);
        b[i + n[0] * (j + n[1] * k)] *= ev;
      }
  b[0] = 0;
  free(lx);
  free(ly);
  free(lz);
  /* inverse z, y, x with normalisation (src/PressureEquation.cpp:167-263) */
  const double nz = (double)(g->Ng[2] - 1) * (g->periodic[2] ? 1.0 : 2.0);
  const double ny = (double)(g->Ng[1] - 1) * (g->periodic[1] ? 1.0 : 2.0);
  const double nx = (double)(g->Ng[0] - 1) * (g->periodic[0] ? 1.0 : 2.0);
  transform_axis(g, b, n, 2, AX_INV, nz);
  transform_axis(g, b, n, 1, AX_INV, ny);
  transform_axis(g, b, n, 0, AX_INV, nx);
}

void orc_poisson_packed_direct(const orc_grid *g, double *b) {
  g_direct_transforms = 1;
  orc_poisson_packed(g, b);
  g_direct_transforms = 0;
}

void orc_set_direct_transforms(int on) { g_direct_transforms = on ? 1 : 0; }

/* ------------------------------------------------------------------------ */
/* Cell-centred Neumann axes (no reference counterpart: the reference's      */
/* Neumann axes are node-centred, FFTW_REDFT00).  N unknowns at the cell      */
/* centres, ghosts phi_{-1} = phi_0 and phi_N = phi_{N-1}; the eigenvectors   */
/* of the 1-D operator are cos(pi k (j + 1/2) / N), eigenvalues               */
/* 2 (cos(pi k / N) - 1) / h^2, so the transform pair is the DCT-II           */
/*   X_k = sum_j x_j cos(pi k (2j + 1) / (2N))                                */
/* and the DCT-III y_j = X_0 + 2 sum_{k>=1} X_k cos(pi k (2j + 1) / (2N)),    */
/* which returns N x.  Direct O(N^2) sums over an exact cosine table.         */
/* ------------------------------------------------------------------------ */
static double *cos_table_4n(long N) {
  double *c = malloc(sizeof(double) * 4 * N);
  for (long m = 0; m < 4 * N; m++) c[m] = cos(M_PI * (double)m / (2.0 * (double)N));
  return c;
}

static void dct2_line(long N, const double *c, const double *x, double *X) {
  for (long k = 0; k < N; k++) {
    double s = 0;
    for (long j = 0; j < N; j++) s += x[j] * c[(k * (2 * j + 1)) % (4 * N)];
    X[k] = s;
  }
}

static void dct3_line(long N, const double *c, const double *X, double *y) {
  for (long j = 0; j < N; j++) {
    double s = X[0];
    for (long k = 1; k < N; k++) s += 2.0 * X[k] * c[(k * (2 * j + 1)) % (4 * N)];
    y[j] = s;
  }
}

void orc_dct2_direct(long n, const double *in, double *out) {
  double *c = cos_table_4n(n);
  dct2_line(n, c, in, out);
  free(c);
}

void orc_dct3_direct(long n, const double *in, double *out) {
  double *c = cos_table_4n(n);
  dct3_line(n, c, in, out);
  free(c);
}

static double eig_cell(long index, double delta, long N) {
  return 2.0 * (cos(M_PI * (double)index / (double)N) - 1.0) / (delta * delta);
}

/* DCT-II (AX_FWD) or DCT-III divided by N (AX_INV) along axis a. */
static void transform_axis_cell(double *b, const long n[3], int a, ax_dir dir) {
  const long len = n[a];
  const long stride = a == 0 ? 1 : (a == 1 ? n[0] : n[0] * n[1]);
  const int b1 = a == 0 ? 1 : 0, b2 = a == 2 ? 1 : 2;
  const long nl1 = n[b1], nl2 = n[b2];
  const long s1 = b1 == 0 ? 1 : n[0], s2 = b2 == 1 ? n[0] : n[0] * n[1];
  double *c = cos_table_4n(len);
#pragma omp parallel
  {
    double *lin = malloc(sizeof(double) * len), *lout = malloc(sizeof(double) * len);
#pragma omp for collapse(2) schedule(static)
    for (long q2 = 0; q2 < nl2; q2++)
      for (long q1 = 0; q1 < nl1; q1++) {
        double *base = b + q1 * s1 + q2 * s2;
        for (long t = 0; t < len; t++) lin[t] = base[t * stride];
        if (dir == AX_FWD) {
          dct2_line(len, c, lin, lout);
          for (long t = 0; t < len; t++) base[t * stride] = lout[t];
        } else {
          dct3_line(len, c, lin, lout);
          for (long t = 0; t < len; t++) base[t * stride] = lout[t] / (double)len;
        }
      }
    free(lin);
    free(lout);
  }
  free(c);
}

void orc_poisson_packed_cell(const orc_grid *g, const int cell[3], double *b) {
  long n[3];
  for (int a = 0; a < 3; a++) n[a] = cell[a] ? g->Ng[a] - 1 : g->npts[a];
  for (int a = 0; a < 3; a++) {
    if (cell[a])
      transform_axis_cell(b, n, a, AX_FWD);
    else
      transform_axis(g, b, n, a, AX_FWD, 1.0);
  }
  double *lam[3];
  for (int a = 0; a < 3; a++) {
    lam[a] = malloc(sizeof(double) * n[a]);
    for (long i = 0; i < n[a]; i++)
      lam[a][i] = cell[a] ? eig_cell(i, g->h[a], n[a])
                          : (g->periodic[a] ? eig_periodic(i, g->h[a], n[a]) : eig_neumann(i, g->h[a], n[a]));
  }
#pragma omp parallel for schedule(static)
  for (long k = 0; k < n[2]; k++)
    for (long j = 0; j < n[1]; j++)
      for (long i = 0; i < n[0]; i++) b[i + n[0] * (j + n[1] * k)] *= 1 / (lam[0][i] + lam[1][j] + lam[2][k]);
  b[0] = 0;
  for (int a = 0; a < 3; a++) free(lam[a]);
  for (int a = 2; a >= 0; a--) {
    if (cell[a])
      transform_axis_cell(b, n, a, AX_INV);
    else
      transform_axis(g, b, n, a, AX_INV, (double)(g->Ng[a] - 1) * (g->periodic[a] ? 1.0 : 2.0));
  }
}

/* Owner-point ranges of the pressure tensor (STAGGERED_TENSOR_ITERATE_OVER_
 * ALL_OWNER_POINTS, include/StaggeredTensorMacros.h:39-85 at one rank). */
static void owner_range(const orc_grid *g, long lo[3], long hi[3]) {
  for (int a = 0; a < 3; a++) {
    lo[a] = g->periodic[a] ? 1 : 0;
    hi[a] = g->periodic[a] ? g->N[a] - 1 : g->N[a];
  }
}

/* apply_non_homogeneous_neumann (src/PressureEquation.cpp:10-57). */
static void apply_nhn(const orc_grid *g, double *rhs, const mif_function *gp) {
  long e[3];
  orc_extents(g, 3, e);
  for (long j = 0; j < e[1]; j++)
    for (long i = 0; i < e[0]; i++)
      AT(rhs, e, i, j, 0) += eval_node(g, gp, 2, 0.0, i, j, 0) * 2.0 * g->one_over_h[2];
  for (long j = 0; j < e[1]; j++)
    for (long i = 0; i < e[0]; i++)
      AT(rhs, e, i, j, e[2] - 1) -= eval_node(g, gp, 2, 0.0, i, j, e[2] - 1) * 2.0 * g->one_over_h[2];
  for (long k = 0; k < e[2]; k++)
    for (long i = 0; i < e[0]; i++)
      AT(rhs, e, i, 0, k) += eval_node(g, gp, 1, 0.0, i, 0, k) * 2.0 * g->one_over_h[1];
  for (long k = 0; k < e[2]; k++)
    for (long i = 0; i < e[0]; i++)
      AT(rhs, e, i, e[1] - 1, k) -= eval_node(g, gp, 1, 0.0, i, e[1] - 1, k) * 2.0 * g->one_over_h[1];
  for (long j = 0; j < e[1]; j++)
    for (long k = 0; k < e[2]; k++)
      AT(rhs, e, 0, j, k) += eval_node(g, gp, 0, 0.0, 0, j, k) * 2.0 * g->one_over_h[0];
  for (long j = 0; j < e[1]; j++)
    for (long k = 0; k < e[2]; k++)
      AT(rhs, e, e[0] - 1, j, k) -= eval_node(g, gp, 0, 0.0, e[0] - 1, j, k) * 2.0 * g->one_over_h[0];
}

/* solve_pressure_equation_homogeneous_periodic / _non_homogeneous_neumann
 * (src/PressureEquation.cpp:266-286): rhs = div/dt on owner points
 * (:59-61), optional Neumann data, pack (src/PressureTensor.cpp:16-19), solve,
 * unpack and periodic copies (src/PressureTensor.cpp:21-34). */
void orc_pressure_rhs(const orc_grid *g, double *p, double *const vel[3], const mif_function *gp, double dt) {
  long e[3], lo[3], hi[3];
  orc_extents(g, 3, e);
  owner_range(g, lo, hi);
#pragma omp parallel for schedule(static)
  for (long k = lo[2]; k < hi[2]; k++)
    for (long j = lo[1]; j < hi[1]; j++)
      for (long i = lo[0]; i < hi[0]; i++) AT(p, e, i, j, k) = divergence(g, vel, i, j, k) / dt;
  if (gp) apply_nhn(g, p, gp);
}

void orc_solve_pressure(const orc_grid *g, double *p, double *const vel[3], const mif_function *gp,
                        double dt) {
  long e[3], lo[3], hi[3];
  orc_extents(g, 3, e);
  owner_range(g, lo, hi);
  orc_pressure_rhs(g, p, vel, gp, dt);
  const long *n = g->npts;
  double *b = malloc(sizeof(double) * (size_t)(n[0] * n[1] * n[2]));
#pragma omp parallel for schedule(static)
  for (long k = lo[2]; k < hi[2]; k++)
    for (long j = lo[1]; j < hi[1]; j++)
      for (long i = lo[0]; i < hi[0]; i++)
        b[(i - lo[0]) + n[0] * ((j - lo[1]) + n[1] * (k - lo[2]))] = AT(p, e, i, j, k);
  orc_poisson_packed(g, b);
#pragma omp parallel for schedule(static)
  for (long k = lo[2]; k < hi[2]; k++)
    for (long j = lo[1]; j < hi[1]; j++)
      for (long i = lo[0]; i < hi[0]; i++)
        AT(p, e, i, j, k) = b[(i - lo[0]) + n[0] * ((j - lo[1]) + n[1] * (k - lo[2]))];
  free(b);
  orc_apply_periodic(g, p, 3);
}

/* ------------------------------------------------------------------------ */
/* RK3 + incremental projection (src/Timestep.cpp:10-156).                   */
/* ------------------------------------------------------------------------ */
static void interior(const orc_grid *g, int s, long e[3]) { orc_extents(g, s, e); }

/* update_velocity_pressure_gradient (src/Timestep.cpp:66-76). */
static void project(const orc_grid *g, double *const vel[3], const double *phi, double dts) {
  for (int c = 0; c < 3; c++) {
    long e[3];
    interior(g, c, e);
    double *f = vel[c];
#pragma omp parallel for schedule(static)
    for (long k = 1; k < e[2] - 1; k++)
      for (long j = 1; j < e[1] - 1; j++)
        for (long i = 1; i < e[0] - 1; i++) AT(f, e, i, j, k) -= pgrad(g, phi, c, i, j, k) * dts;
  }
}

/* update_pressure (src/Timestep.cpp:79-81): all points. */
static void add_pressure(const orc_grid *g, double *p, const double *phi) {
  const long n = orc_field_size(g, 3);
#pragma omp parallel for schedule(static)
  for (long q = 0; q < n; q++) p[q] += phi[q];
}

/* The boundary function timestep_nhn hands to the pressure solve of stage s:
 * PRESSURE_EQUATION_true(velocity, dt, new_time, prev_time) passes
 * exact_pressure_gradient.get_difference_over_time(new_time, prev_time)
 * (src/Timestep.cpp:89-95) with (time_1, t_n), (time_2, time_1) and
 * (final_time, time_2) (:110, :122, :134), and get_difference_over_time(a, b)
 * returns f(b) - f(a) (src/VectorFunction.cpp:53-61): the function is
 * gradp(prev_time) - gradp(new_time). */
void orc_nhn_stage_function(const orc_grid *g, const mif_function *gradp, double t_n, int stage,
                            mif_function *out) {
  const double dt = g->dt;
  const double time_1 = t_n + 64.0 / 120.0 * dt;
  const double time_2 = time_1 + 16.0 / 120.0 * dt;
  const double final_time = t_n + dt;
  memset(out, 0, sizeof(*out));
  out->kind = MIF_FN_TIME_DIFFERENCE;
  out->base = gradp->kind;
  out->Re = gradp->Re;
  memcpy(out->c, gradp->c, sizeof(out->c));
  /* value = base(t1) - base(t0) */
  out->t0 = stage == 1 ? time_1 : (stage == 2 ? time_2 : final_time);
  out->t1 = stage == 1 ? t_n : (stage == 2 ? time_1 : time_2);
}

void orc_timestep(const orc_grid *g, double *const vel[3], double *const buf[3], double *const buf2[3],
                  const mif_function *ev, const mif_function *gradp, double t_n, double *p,
                  double *p_buf) {
  const double dt = g->dt;
  const double dt_1 = 64.0 / 120.0 * dt;
  const double time_1 = t_n + dt_1;
  const double dt_2 = 16.0 / 120.0 * dt;
  const double time_2 = time_1 + dt_2;
  const double dt_3 = 40.0 / 120.0 * dt;
  const double final_time = t_n + dt;
  mif_function gdiff;

  /* Stage 1: COMPUTE_COMPONENT_Y2 (src/Timestep.cpp:10-22) */
  {
    const double a1 = 64.0 / 120.0 * dt, b = a1;
    for (int c = 0; c < 3; c++) {
      long e[3];
      interior(g, c, e);
#pragma omp parallel for schedule(static)
      for (long k = 1; k < e[2] - 1; k++)
        for (long j = 1; j < e[1] - 1; j++)
          for (long i = 1; i < e[0] - 1; i++) {
            const double rhs = orc_momentum_rhs(g, vel, c, i, j, k) + g->force[c];
            const double pg = pgrad(g, p, c, i, j, k);
            AT(buf[c], e, i, j, k) = AT(vel[c], e, i, j, k) + a1 * rhs - b * pg;
            AT(buf2[c], e, i, j, k) = rhs;
          }
    }
  }
  orc_apply_bc(g, buf[0], buf[1], buf[2], ev, time_1);
  if (gradp) orc_nhn_stage_function(g, gradp, t_n, 1, &gdiff);
  orc_solve_pressure(g, p_buf, buf, gradp ? &gdiff : NULL, dt_1);
  add_pressure(g, p, p_buf);
  project(g, buf, p_buf, dt_1);

  /* Stage 2: COMPUTE_COMPONENT_Y3 (src/Timestep.cpp:25-39) */
  {
    const double a1 = -34.0 / 120.0 * dt, a2 = 50.0 / 120.0 * dt, b = a1 + a2;
    for (int c = 0; c < 3; c++) {
      long e[3];
      interior(g, c, e);
#pragma omp parallel for schedule(static)
      for (long k = 1; k < e[2] - 1; k++)
        for (long j = 1; j < e[1] - 1; j++)
          for (long i = 1; i < e[0] - 1; i++) {
            const double rhs_1 = AT(buf2[c], e, i, j, k);
            const double rhs_2 = orc_momentum_rhs(g, buf, c, i, j, k) + g->force[c];
            const double rhs_2_scaled = a2 * rhs_2;
            const double pg = pgrad(g, p, c, i, j, k);
            AT(buf2[c], e, i, j, k) = AT(buf[c], e, i, j, k) + a1 * rhs_1 + rhs_2_scaled - b * pg;
            AT(vel[c], e, i, j, k) = rhs_2_scaled;
          }
    }
  }
  orc_apply_bc(g, buf2[0], buf2[1], buf2[2], ev, time_2);
  if (gradp) orc_nhn_stage_function(g, gradp, t_n, 2, &gdiff);
  orc_solve_pressure(g, p_buf, buf2, gradp ? &gdiff : NULL, dt_2);
  add_pressure(g, p, p_buf);
  project(g, buf2, p_buf, dt_2);

  /* Stage 3: COMPUTE_COMPONENT_U_STAR (src/Timestep.cpp:42-54) */
  {
    const double a2 = -50.0 / 120.0 * dt, a3 = 90.0 / 120.0 * dt, b = a2 + a3;
    for (int c = 0; c < 3; c++) {
      long e[3];
      interior(g, c, e);
#pragma omp parallel for schedule(static)
      for (long k = 1; k < e[2] - 1; k++)
        for (long j = 1; j < e[1] - 1; j++)
          for (long i = 1; i < e[0] - 1; i++) {
            const double rhs_2_scaled = -AT(vel[c], e, i, j, k);
            const double rhs_3 = orc_momentum_rhs(g, buf2, c, i, j, k) + g->force[c];
            const double pg = pgrad(g, p, c, i, j, k);
            AT(vel[c], e, i, j, k) = AT(buf2[c], e, i, j, k) + rhs_2_scaled + a3 * rhs_3 - b * pg;
          }
    }
  }
  orc_apply_bc(g, vel[0], vel[1], vel[2], ev, final_time);
  if (gradp) orc_nhn_stage_function(g, gradp, t_n, 3, &gdiff);
  orc_solve_pressure(g, p_buf, vel, gradp ? &gdiff : NULL, dt_3);
  add_pressure(g, p, p_buf);
  project(g, vel, p_buf, dt_3);
}

/* ------------------------------------------------------------------------ */
/* Velocity-only RK3 (src/TimestepVelocity.cpp:11-90).                       */
/* ------------------------------------------------------------------------ */
void orc_timestep_velocity(const orc_grid *g, double *const vel[3], double *const buf[3],
                           double *const rhsb[3], const mif_function *ev, const mif_function *fo,
                           double t_n) {
  const double c2 = 8.0 / 15.0, a21 = 8.0 / 15.0, c3 = 2.0 / 3.0, a31 = 1.0 / 4.0, a32 = 5.0 / 12.0,
               b1 = 1.0 / 4.0, b3 = 3.0 / 4.0;
  const double dt = g->dt;
  const double time_1 = t_n + c2 * dt;
  const double time_2 = t_n + c3 * dt;
  const double final_time = t_n + dt;
  /* Y2 (src/TimestepVelocity.cpp:20-29) */
  for (int c = 0; c < 3; c++) {
    long e[3];
    interior(g, c, e);
#pragma omp parallel for schedule(static)
    for (long k = 1; k < e[2] - 1; k++)
      for (long j = 1; j < e[1] - 1; j++)
        for (long i = 1; i < e[0] - 1; i++) {
          const double rhs = orc_momentum_rhs(g, vel, c, i, j, k) + eval_at(g, c, fo, c, t_n, i, j, k);
          AT(rhsb[c], e, i, j, k) = rhs;
          AT(buf[c], e, i, j, k) = AT(vel[c], e, i, j, k) + dt * a21 * rhs;
        }
  }
  orc_apply_bc(g, buf[0], buf[1], buf[2], ev, time_1);
  /* Y3 (src/TimestepVelocity.cpp:32-41) */
  for (int c = 0; c < 3; c++) {
    long e[3];
    interior(g, c, e);
#pragma omp parallel for schedule(static)
    for (long k = 1; k < e[2] - 1; k++)
      for (long j = 1; j < e[1] - 1; j++)
        for (long i = 1; i < e[0] - 1; i++) {
          const double rhs = AT(rhsb[c], e, i, j, k);
          AT(rhsb[c], e, i, j, k) = AT(vel[c], e, i, j, k) + dt * (b1 * rhs);
          const double f2 = orc_momentum_rhs(g, buf, c, i, j, k) + eval_at(g, c, fo, c, time_1, i, j, k);
          AT(vel[c], e, i, j, k) = AT(vel[c], e, i, j, k) + dt * (a31 * rhs + a32 * f2);
        }
  }
  orc_apply_bc(g, vel[0], vel[1], vel[2], ev, time_2);
  /* U* (src/TimestepVelocity.cpp:44-51) */
  for (int c = 0; c < 3; c++) {
    long e[3];
    interior(g, c, e);
#pragma omp parallel for schedule(static)
    for (long k = 1; k < e[2] - 1; k++)
      for (long j = 1; j < e[1] - 1; j++)
        for (long i = 1; i < e[0] - 1; i++) {
          const double rhs = AT(rhsb[c], e, i, j, k);
          const double f3 = orc_momentum_rhs(g, vel, c, i, j, k) + eval_at(g, c, fo, c, time_2, i, j, k);
          AT(buf[c], e, i, j, k) = rhs + dt * (b3 * f3);
        }
  }
  orc_apply_bc(g, buf[0], buf[1], buf[2], ev, final_time);
  /* velocity.swap_data(velocity_buffer) (src/TimestepVelocity.cpp:89) */
  for (int c = 0; c < 3; c++) {
    const long n = orc_field_size(g, c);
    double *tmp = malloc(sizeof(double) * n);
    memcpy(tmp, vel[c], sizeof(double) * n);
    memcpy(vel[c], buf[c], sizeof(double) * n);
    memcpy(buf[c], tmp, sizeof(double) * n);
    free(tmp);
  }
}

/* ------------------------------------------------------------------------ */
/* adjust_pressure and error norms.                                          */
/* ------------------------------------------------------------------------ */
void orc_adjust_pressure(const orc_grid *g, double *p, const mif_function *fn, double t) {
  long e[3], lo[3], hi[3];
  orc_extents(g, 3, e);
  owner_range(g, lo, hi);
  double diff = 0;
  for (long k = lo[2]; k < hi[2]; k++)
    for (long j = lo[1]; j < hi[1]; j++)
      for (long i = lo[0]; i < hi[0]; i++) diff += eval_node(g, fn, 0, t, i, j, k) - AT(p, e, i, j, k);
  const long nx = g->npts[0], ny = g->npts[1], nz = g->npts[2];
  diff /= (double)(nx * ny * nz);
  const long n = e[0] * e[1] * e[2];
  for (long q = 0; q < n; q++) p[q] += diff;
}

void orc_error_norms_velocity(const orc_grid *g, double *const vel[3], const mif_function *fn, double t,
                              double out[3]) {
  long eu[3], ev[3], ew[3];
  orc_extents(g, 0, eu);
  orc_extents(g, 1, ev);
  orc_extents(g, 2, ew);
  double l1 = 0, l2 = 0, linf = 0;
  for (long k = 1; k < g->N[2] - 1; k++) {
    const double z = g->min[2] + (double)(g->base[2] + k) * g->h[2];
    for (long j = 1; j < g->N[1] - 1; j++) {
      const double y = g->min[1] + (double)(g->base[1] + j) * g->h[1];
      for (long i = 1; i < g->N[0] - 1; i++) {
        const double x = g->min[0] + (double)(g->base[0] + i) * g->h[0];
        const double iu = (AT(vel[0], eu, i, j, k) + AT(vel[0], eu, i + 1, j, k)) / 2.0;
        const double iv = (AT(vel[1], ev, i, j, k) + AT(vel[1], ev, i, j + 1, k)) / 2.0;
        const double iw = (AT(vel[2], ew, i, j, k) + AT(vel[2], ew, i, j, k + 1)) / 2.0;
        const double eu_ = orc_fn_eval(fn, 0, t, x, y, z) - iu;
        const double ev_ = orc_fn_eval(fn, 1, t, x, y, z) - iv;
        const double ew_ = orc_fn_eval(fn, 2, t, x, y, z) - iw;
        l1 += sqrt(eu_ * eu_ + ev_ * ev_ + ew_ * ew_);
        l2 += eu_ * eu_ + ev_ * ev_ + ew_ * ew_;
        double m = fabs(eu_);
        if (fabs(ev_) > m) m = fabs(ev_);
        if (fabs(ew_) > m) m = fabs(ew_);
        if (m > linf) linf = m;
      }
    }
  }
  out[0] = l1 * g->h[0] * g->h[1] * g->h[2];
  out[1] = sqrt(l2 * g->h[0] * g->h[1] * g->h[2]);
  out[2] = linf;
}

void orc_error_norms_pressure(const orc_grid *g, const double *p, const mif_function *fn, double t,
                              double out[3]) {
  long e[3], lo[3], hi[3];
  orc_extents(g, 3, e);
  owner_range(g, lo, hi);
  double l1 = 0, l2 = 0, linf = 0;
  for (long k = lo[2]; k < hi[2]; k++) {
    const double z = g->min[2] + (double)(g->base[2] + k) * g->h[2];
    for (long j = lo[1]; j < hi[1]; j++) {
      const double y = g->min[1] + (double)(g->base[1] + j) * g->h[1];
      for (long i = lo[0]; i < hi[0]; i++) {
        const double x = g->min[0] + (double)(g->base[0] + i) * g->h[0];
        const double err = orc_fn_eval(fn, 0, t, x, y, z) - AT(p, e, i, j, k);
        l1 += fabs(err);
        l2 += err * err;
        if (fabs(err) > linf) linf = fabs(err);
      }
    }
  }
  out[0] = l1 * g->h[0] * g->h[1] * g->h[2];
  out[1] = sqrt(l2 * g->h[0] * g->h[1] * g->h[2]);
  out[2] = linf;
}

/* Pressure-gradient field on interior points (VELOCITY_TENSOR_SET_FOR_ALL_
 * POINTS with pressure_gradient_{u,v,w}, test/full_test.cpp:141-142). */
void orc_pressure_gradient(const orc_grid *g, const double *p, double *const out[3]) {
  for (int c = 0; c < 3; c++) {
    long e[3];
    interior(g, c, e);
    for (long k = 1; k < e[2] - 1; k++)
      for (long j = 1; j < e[1] - 1; j++)
        for (long i = 1; i < e[0] - 1; i++) AT(out[c], e, i, j, k) = pgrad(g, p, c, i, j, k);
  }
}

/* ------------------------------------------------------------------------ */
/* Adjoint of one timestep() (no reference counterpart).                     */
/*                                                                          */
/* The step is a composition of: the stage updates (N(X) = momentum RHS,    */
/* quadratic in X), the boundary map (Dirichlet faces = constants, periodic */
/* ghosts = copies of interior points), the divergence, the spectral solve  */
/* M, the unpack with periodic images, p += phi and the interior projection */
/* u -= dt_s G phi.  Each linear piece is transposed as a scatter here; the */
/* spectral solve's transpose is W M W^-1 with W the trapezoid weights (the */
/* DCT-I matrix is diag(c) times a symmetric matrix, c = 1 at the two ends  */
/* and 2 inside).  Serial scatters keep the oracle simple and deterministic. */
/* ------------------------------------------------------------------------ */
static int interior_pt(const orc_grid *g, int c, long i, long j, long k) {
  long e[3];
  orc_extents(g, c, e);
  return i >= 1 && j >= 1 && k >= 1 && i <= e[0] - 2 && j <= e[1] - 2 && k <= e[2] - 2;
}

static double *vzeros(const orc_grid *g, int s) { return (double *)calloc((size_t)orc_field_size(g, s), sizeof(double)); }

/* Xb += J_N(X)^T Nb (Nb holds values at interior points of each component). */
static void adj_convdiff(const orc_grid *g, double *const X[3], double *const Nb[3], double *const Xb[3]) {
  long e[3][3];
  for (int c = 0; c < 3; c++) orc_extents(g, c, e[c]);
  const double *u = X[0], *v = X[1], *w = X[2];
  double *ub = Xb[0], *vb = Xb[1], *wb = Xb[2];
  const double Dx = g->one_over_h2Re[0], Dy = g->one_over_h2Re[1], Dz = g->one_over_h2Re[2];
  const double A2 = g->one_over_2h[0], B2 = g->one_over_2h[1], C2 = g->one_over_2h[2];
  const double A8 = g->one_over_8h[0], B8 = g->one_over_8h[1], C8 = g->one_over_8h[2];
#define U(i, j, k) AT(u, e[0], i, j, k)
#define V(i, j, k) AT(v, e[1], i, j, k)
#define W(i, j, k) AT(w, e[2], i, j, k)
#define UB(i, j, k) AT(ub, e[0], i, j, k)
#define VB(i, j, k) AT(vb, e[1], i, j, k)
#define WB(i, j, k) AT(wb, e[2], i, j, k)
  /* u component */
  for (long k = 1; k < e[0][2] - 1; k++)
    for (long j = 1; j < e[0][1] - 1; j++)
      for (long i = 1; i < e[0][0] - 1; i++) {
        const double n = AT(Nb[0], e[0], i, j, k);
        if (n == 0.0) continue;
        const double dux = U(i + 1, j, k) - U(i - 1, j, k), duy = U(i, j + 1, k) - U(i, j - 1, k),
                     duz = U(i, j, k + 1) - U(i, j, k - 1);
        const double Vs = V(i, j + 1, k) + V(i, j, k) + V(i - 1, j + 1, k) + V(i - 1, j, k);
        const double Ws = W(i, j, k + 1) + W(i, j, k) + W(i - 1, j, k + 1) + W(i - 1, j, k);
        UB(i, j, k) += n * (-dux * A2 - 2 * (Dx + Dy + Dz));
        UB(i + 1, j, k) += n * (-U(i, j, k) * A2 + Dx);
        UB(i - 1, j, k) += n * (U(i, j, k) * A2 + Dx);
        UB(i, j + 1, k) += n * (-Vs * B8 + Dy);
        UB(i, j - 1, k) += n * (Vs * B8 + Dy);
        UB(i, j, k + 1) += n * (-Ws * C8 + Dz);
        UB(i, j, k - 1) += n * (Ws * C8 + Dz);
        const double cv = -duy * B8, cw = -duz * C8;
        VB(i, j + 1, k) += n * cv;
        VB(i, j, k) += n * cv;
        VB(i - 1, j + 1, k) += n * cv;
        VB(i - 1, j, k) += n * cv;
        WB(i, j, k + 1) += n * cw;
        WB(i, j, k) += n * cw;
        WB(i - 1, j, k + 1) += n * cw;
        WB(i - 1, j, k) += n * cw;
      }
  /* v component */
  for (long k = 1; k < e[1][2] - 1; k++)
    for (long j = 1; j < e[1][1] - 1; j++)
      for (long i = 1; i < e[1][0] - 1; i++) {
        const double n = AT(Nb[1], e[1], i, j, k);
        if (n == 0.0) continue;
        const double dvx = V(i + 1, j, k) - V(i - 1, j, k), dvy = V(i, j + 1, k) - V(i, j - 1, k),
                     dvz = V(i, j, k + 1) - V(i, j, k - 1);
        const double Us = U(i + 1, j, k) + U(i, j, k) + U(i + 1, j - 1, k) + U(i, j - 1, k);
        const double Ws = W(i, j, k + 1) + W(i, j, k) + W(i, j - 1, k + 1) + W(i, j - 1, k);
        VB(i, j, k) += n * (-dvy * B2 - 2 * (Dx + Dy + Dz));
        VB(i + 1, j, k) += n * (-Us * A8 + Dx);
        VB(i - 1, j, k) += n * (Us * A8 + Dx);
        VB(i, j + 1, k) += n * (-V(i, j, k) * B2 + Dy);
        VB(i, j - 1, k) += n * (V(i, j, k) * B2 + Dy);
        VB(i, j, k + 1) += n * (-Ws * C8 + Dz);
        VB(i, j, k - 1) += n * (Ws * C8 + Dz);
        const double cu = -dvx * A8, cw = -dvz * C8;
        UB(i + 1, j, k) += n * cu;
        UB(i, j, k) += n * cu;
        UB(i + 1, j - 1, k) += n * cu;
        UB(i, j - 1, k) += n * cu;
        WB(i, j, k + 1) += n * cw;
        WB(i, j, k) += n * cw;
        WB(i, j - 1, k + 1) += n * cw;
        WB(i, j - 1, k) += n * cw;
      }
  /* w component */
  for (long k = 1; k < e[2][2] - 1; k++)
    for (long j = 1; j < e[2][1] - 1; j++)
      for (long i = 1; i < e[2][0] - 1; i++) {
        const double n = AT(Nb[2], e[2], i, j, k);
        if (n == 0.0) continue;
        const double dwx = W(i + 1, j, k) - W(i - 1, j, k), dwy = W(i, j + 1, k) - W(i, j - 1, k),
                     dwz = W(i, j, k + 1) - W(i, j, k - 1);
        const double Us = U(i + 1, j, k) + U(i, j, k) + U(i + 1, j, k - 1) + U(i, j, k - 1);
        const double Vs = V(i, j + 1, k) + V(i, j, k) + V(i, j + 1, k - 1) + V(i, j, k - 1);
        WB(i, j, k) += n * (-dwz * C2 - 2 * (Dx + Dy + Dz));
        WB(i + 1, j, k) += n * (-Us * A8 + Dx);
        WB(i - 1, j, k) += n * (Us * A8 + Dx);
        WB(i, j + 1, k) += n * (-Vs * B8 + Dy);
        WB(i, j - 1, k) += n * (Vs * B8 + Dy);
        WB(i, j, k + 1) += n * (-W(i, j, k) * C2 + Dz);
        WB(i, j, k - 1) += n * (W(i, j, k) * C2 + Dz);
        const double cu = -dwx * A8, cv = -dwy * B8;
        UB(i + 1, j, k) += n * cu;
        UB(i, j, k) += n * cu;
        UB(i + 1, j, k - 1) += n * cu;
        UB(i, j, k - 1) += n * cu;
        VB(i, j + 1, k) += n * cv;
        VB(i, j, k) += n * cv;
        VB(i, j + 1, k - 1) += n * cv;
        VB(i, j, k - 1) += n * cv;
      }
#undef U
#undef V
#undef W
#undef UB
#undef VB
#undef WB
}

/* pb += G^T applied to Ob with coefficient b: the forward term of component
 * c at interior Q is -b (p(Q) - p(Q - e_c)) / h_c. */
static void adj_grad(const orc_grid *g, double *const Ob[3], double b, double *pb) {
  long ep[3];
  orc_extents(g, 3, ep);
  const long off[3] = {1, ep[0], ep[0] * ep[1]};
  for (int c = 0; c < 3; c++) {
    long e[3];
    orc_extents(g, c, e);
    for (long k = 1; k < e[2] - 1; k++)
      for (long j = 1; j < e[1] - 1; j++)
        for (long i = 1; i < e[0] - 1; i++) {
          const double o = AT(Ob[c], e, i, j, k) * b * g->one_over_h[c];
          const long q = i + ep[0] * (j + ep[1] * k);
          pb[q] -= o;
          pb[q - off[c]] += o;
        }
  }
}

/* Source owner index of a point of the pressure extents (periodic map). */
static long owner_of(const orc_grid *g, long i, long j, long k) {
  long s[3] = {i, j, k};
  for (int a = 0; a < 3; a++) {
    if (!g->periodic[a]) continue;
    if (s[a] == 0) s[a] = g->N[a] - 2;
    else if (s[a] == g->N[a] - 1) s[a] = 1;
    s[a] -= 1;
  }
  return s[0] + g->npts[0] * (s[1] + g->npts[1] * s[2]);
}

static double trap_w(const orc_grid *g, int a, long i) {
  if (g->periodic[a]) return 1.0;
  return (i == 0 || i == g->npts[a] - 1) ? 0.5 : 1.0;
}

/* yb += (D^T P^T M^T U^T phib) / dts, M^T = W M W^-1. */
static void adj_solve(const orc_grid *g, const double *phib, double dts, double *const yb[3]) {
  long ep[3];
  orc_extents(g, 3, ep);
  const long *n = g->npts;
  double *b = (double *)calloc((size_t)(n[0] * n[1] * n[2]), sizeof(double));
  for (long k = 0; k < ep[2]; k++)
    for (long j = 0; j < ep[1]; j++)
      for (long i = 0; i < ep[0]; i++) b[owner_of(g, i, j, k)] += AT(phib, ep, i, j, k);
  for (long k = 0; k < n[2]; k++)
    for (long j = 0; j < n[1]; j++)
      for (long i = 0; i < n[0]; i++)
        b[i + n[0] * (j + n[1] * k)] /= trap_w(g, 0, i) * trap_w(g, 1, j) * trap_w(g, 2, k);
  orc_poisson_packed(g, b);
  for (long k = 0; k < n[2]; k++)
    for (long j = 0; j < n[1]; j++)
      for (long i = 0; i < n[0]; i++)
        b[i + n[0] * (j + n[1] * k)] *= trap_w(g, 0, i) * trap_w(g, 1, j) * trap_w(g, 2, k) / dts;
  long lo[3];
  for (int a = 0; a < 3; a++) lo[a] = g->periodic[a] ? 1 : 0;
  long e[3][3];
  for (int c = 0; c < 3; c++) orc_extents(g, c, e[c]);
  for (long k = 0; k < n[2]; k++)
    for (long j = 0; j < n[1]; j++)
      for (long i = 0; i < n[0]; i++) {
        const double bb = b[i + n[0] * (j + n[1] * k)];
        const long I = i + lo[0], J = j + lo[1], K = k + lo[2];
        AT(yb[0], e[0], I + 1, J, K) += bb * g->one_over_h[0];
        AT(yb[0], e[0], I, J, K) -= bb * g->one_over_h[0];
        AT(yb[1], e[1], I, J + 1, K) += bb * g->one_over_h[1];
        AT(yb[1], e[1], I, J, K) -= bb * g->one_over_h[1];
        AT(yb[2], e[2], I, J, K + 1) += bb * g->one_over_h[2];
        AT(yb[2], e[2], I, J, K) -= bb * g->one_over_h[2];
      }
  free(b);
}

/* Yb = B^T yb: an interior point collects its own adjoint and those of the
 * periodic ghosts copied from it; boundary-shell entries become zero
 * (Dirichlet values are constants of the map). */
static void adj_bc(const orc_grid *g, double *const yb[3], double *const Yb[3]) {
  for (int c = 0; c < 3; c++) {
    long e[3];
    orc_extents(g, c, e);
    memset(Yb[c], 0, sizeof(double) * (size_t)orc_field_size(g, c));
    for (long k = 1; k < e[2] - 1; k++)
      for (long j = 1; j < e[1] - 1; j++)
        for (long i = 1; i < e[0] - 1; i++) {
          const long P[3] = {i, j, k};
          long opts[3][3];
          int nopt[3];
          for (int a = 0; a < 3; a++) {
            nopt[a] = 1;
            opts[a][0] = P[a];
            if (g->periodic[a]) {
              if (P[a] == 1) opts[a][nopt[a]++] = e[a] - 1;
              if (P[a] == e[a] - 2) opts[a][nopt[a]++] = 0;
            }
          }
          double s = 0;
          for (int qa = 0; qa < nopt[0]; qa++)
            for (int qb = 0; qb < nopt[1]; qb++)
              for (int qc = 0; qc < nopt[2]; qc++) s += AT(yb[c], e, opts[0][qa], opts[1][qb], opts[2][qc]);
          AT(Yb[c], e, i, j, k) = s;
        }
  }
}

static void interior_scaled(const orc_grid *g, double *const src[3], double a, double *const dst[3]) {
  for (int c = 0; c < 3; c++) {
    long e[3];
    orc_extents(g, c, e);
    memset(dst[c], 0, sizeof(double) * (size_t)orc_field_size(g, c));
    for (long k = 1; k < e[2] - 1; k++)
      for (long j = 1; j < e[1] - 1; j++)
        for (long i = 1; i < e[0] - 1; i++) AT(dst[c], e, i, j, k) = a * AT(src[c], e, i, j, k);
  }
}

static void add_force_bar(const orc_grid *g, double *const Nb[3], double fbar[3]) {
  for (int c = 0; c < 3; c++) {
    long e[3];
    orc_extents(g, c, e);
    double s = 0;
    for (long k = 1; k < e[2] - 1; k++)
      for (long j = 1; j < e[1] - 1; j++)
        for (long i = 1; i < e[0] - 1; i++) s += AT(Nb[c], e, i, j, k);
    fbar[c] += s;
  }
}

void orc_timestep_adjoint(const orc_grid *g, double *const un[3], double *const y2[3], double *const y3[3],
                          double *const ubar[3], double *pbar, double fbar[3]) {
  const double dt = g->dt;
  const double dt_1 = 64.0 / 120.0 * dt, dt_2 = 16.0 / 120.0 * dt, dt_3 = 40.0 / 120.0 * dt;
  const double s1_a1 = 64.0 / 120.0 * dt, s1_b = s1_a1;
  const double s2_a1 = -34.0 / 120.0 * dt, s2_a2 = 50.0 / 120.0 * dt, s2_b = s2_a1 + s2_a2;
  const double s3_a2 = -50.0 / 120.0 * dt, s3_a3 = 90.0 / 120.0 * dt, s3_b = s3_a2 + s3_a3;
  double *t1[3], *t2[3], *Sb[3], *R1b[3], *Nb[3];
  for (int c = 0; c < 3; c++) {
    t1[c] = vzeros(g, c);
    t2[c] = vzeros(g, c);
    Sb[c] = vzeros(g, c);
    R1b[c] = vzeros(g, c);
    Nb[c] = vzeros(g, c);
  }
  const long np = orc_field_size(g, 3);
  double *phib = (double *)malloc(sizeof(double) * (size_t)np);

  /* stage 3: u^{n+1} = B(u') - dt_3 G phi_3, p^{n+1} = p_2 + phi_3 */
  memcpy(phib, pbar, sizeof(double) * (size_t)np);
  adj_grad(g, ubar, dt_3, phib);
  adj_solve(g, phib, dt_3, ubar);
  adj_bc(g, ubar, t1); /* t1 = adjoint of u' (interior) */
  interior_scaled(g, t1, s3_a3, Nb);
  interior_scaled(g, t1, 1.0, t2);
  adj_convdiff(g, y3, Nb, t2); /* t2 = adjoint of Y3' */
  interior_scaled(g, t1, -1.0, Sb);
  adj_grad(g, t1, s3_b, pbar); /* pbar = adjoint of p_2 */
  add_force_bar(g, Nb, fbar);

  /* stage 2: Y3 = Y2' + a1 R1 + a2 N(Y2'), S = a2 N(Y2') */
  memcpy(phib, pbar, sizeof(double) * (size_t)np);
  adj_grad(g, t2, dt_2, phib);
  adj_solve(g, phib, dt_2, t2);
  adj_bc(g, t2, t1); /* t1 = adjoint of Y3 (interior) */
  for (int c = 0; c < 3; c++) {
    const long n = orc_field_size(g, c);
    for (long q = 0; q < n; q++) Nb[c][q] = s2_a2 * (t1[c][q] + Sb[c][q]);
  }
  interior_scaled(g, t1, 1.0, t2);
  adj_convdiff(g, y2, Nb, t2); /* t2 = adjoint of Y2' */
  interior_scaled(g, t1, s2_a1, R1b);
  adj_grad(g, t1, s2_b, pbar); /* pbar = adjoint of p_1 */
  add_force_bar(g, Nb, fbar);

  /* stage 1: Y2 = u^n + a1 N(u^n) - b G p, R1 = N(u^n) */
  memcpy(phib, pbar, sizeof(double) * (size_t)np);
  adj_grad(g, t2, dt_1, phib);
  adj_solve(g, phib, dt_1, t2);
  adj_bc(g, t2, t1); /* t1 = adjoint of Y2 (interior) */
  for (int c = 0; c < 3; c++) {
    const long n = orc_field_size(g, c);
    for (long q = 0; q < n; q++) Nb[c][q] = s1_a1 * t1[c][q] + R1b[c][q];
  }
  interior_scaled(g, t1, 1.0, ubar);
  adj_convdiff(g, un, Nb, ubar); /* ubar = adjoint of u^n */
  adj_grad(g, t1, s1_b, pbar);   /* pbar = adjoint of p^n */
  add_force_bar(g, Nb, fbar);

  for (int c = 0; c < 3; c++) {
    free(t1[c]);
    free(t2[c]);
    free(Sb[c]);
    free(R1b[c]);
    free(Nb[c]);
  }
  free(phib);
  (void)interior_pt;
}
