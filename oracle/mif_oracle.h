/*
 * mif_oracle.h -- CPU oracle for the incompressible Navier-Stokes timestep.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * solver's single-rank (Py = Pz = 1) hot path, written from the reference's
 * algorithm, used by tests/ and by bench.py's CPU-baseline leg to check and
 * time the HIP library.  Nothing in the product links or calls it.
 *
 * Fields are stored exactly as the reference stores a Tensor<Real,3>
 * (include/Tensor.h:68,219-238): flat, index i + sx*(j + sy*k), x fastest,
 * with the per-staggering extents of src/StaggeredTensor.cpp:5-9 and
 * src/Constants.cpp:80-85 at one rank.
 *
 * Pinning (what checks each part; see tests/test_oracle_*.py):
 *   transforms      scipy DCT-I / numpy rfft fixtures, O(N^2) direct sums,
 *                   eigenvector identities of the 1-D operators
 *   Poisson solve   residual of the exact 7-point operator, dense
 *                   pseudo-inverse solves on tiny grids, manufactured field
 *   stencils        exactness on polynomials, O(h^2) truncation
 *   RK stages       stability polynomial on a linear problem
 *   BC              closed-form face values, face order at edges
 *   timestep        manufactured-solution errors and second-order convergence
 *   norms, adjust   closed forms
 */
#ifndef MIF_ORACLE_H
#define MIF_ORACLE_H

#include "../include/mif_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Derived single-rank constants (src/Constants.cpp:58-120 at Py = Pz = 1). */
typedef struct orc_grid {
  long Ng[3];      /* Nx_global, Ny_global, Nz_global                       */
  int periodic[3];
  double L[3], min[3];
  double Re, dt;
  long N[3];       /* Nx, Ny, Nz: node extents incl. periodic ghosts         */
  long base[3];    /* base_i, base_j, base_k                                 */
  double h[3];     /* dx, dy, dz                                             */
  double one_over_2h[3], one_over_8h[3], one_over_h2Re[3], h_over_2[3],
      one_over_h[3];
  long npts[3];    /* Poisson unknowns per axis (Nx_points...)              */
  double force[3]; /* constant body force added to the momentum right-hand
                      side of timestep() (no reference counterpart; zero
                      reproduces the reference)                           */
} orc_grid;

int orc_grid_init(orc_grid *g, const mif_grid *desc);
/* Extents of a field with staggering s (0=x,1=y,2=z,3=none). */
void orc_extents(const orc_grid *g, int s, long ext[3]);
long orc_field_size(const orc_grid *g, int s);

/* Analytic function evaluation (closed forms of the reference's generated
 * functions). */
double orc_fn_eval(const mif_function *f, int comp, double t, double x,
                   double y, double z);

/* StaggeredTensor::set / VelocityTensor::set at time t
 * (src/StaggeredTensor.cpp:205-219). */
void orc_set(const orc_grid *g, double *f, int s, const mif_function *fn,
             int comp, double t, int include_border);

/* VelocityTensor::apply_bc (src/VelocityTensor.cpp:36-233). */
void orc_apply_bc(const orc_grid *g, double *u, double *v, double *w,
                  const mif_function *fn, double t);
/* StaggeredTensor::apply_periodic_bc (src/StaggeredTensor.cpp:221-256). */
void orc_apply_periodic(const orc_grid *g, double *f, int s);

/* Full RK3 + projection step (src/Timestep.cpp:97-156).  vel, buf, buf2 are
 * arrays of 3 component pointers; grad_p is NULL for timestep() and the exact
 * pressure gradient for timestep_nhn(). */
void orc_timestep(const orc_grid *g, double *const vel[3], double *const buf[3],
                  double *const buf2[3], const mif_function *exact_velocity,
                  const mif_function *grad_p, double t_n, double *p,
                  double *p_buf);

/* The pressure-gradient boundary function that timestep_nhn passes to the
 * pressure solve of stage 1, 2 or 3 (src/Timestep.cpp:89-95,110,122,134). */
void orc_nhn_stage_function(const orc_grid *g, const mif_function *gradp, double t_n, int stage,
                            mif_function *out);

/* Velocity-only RK3 with forcing (src/TimestepVelocity.cpp:60-90).  On return
 * vel holds the new velocity (the reference swaps buffers; the oracle copies). */
void orc_timestep_velocity(const orc_grid *g, double *const vel[3],
                           double *const buf[3], double *const rhs[3],
                           const mif_function *exact_velocity,
                           const mif_function *forcing, double t_n);

/* solve_pressure_equation_homogeneous_periodic / _non_homogeneous_neumann
 * (src/PressureEquation.cpp:266-286): p receives the solution (with periodic
 * ghosts).  grad_p NULL selects the homogeneous form. */
void orc_solve_pressure(const orc_grid *g, double *p, double *const vel[3],
                        const mif_function *grad_p, double dt);

/* The right-hand side alone: div(vel)/dt on the owner points of p
 * (compute_rhs_homogeneous_periodic, src/PressureEquation.cpp:59-61), plus
 * the Neumann data of grad_p when it is not NULL
 * (apply_non_homogeneous_neumann, :10-57); other points are left untouched. */
void orc_pressure_rhs(const orc_grid *g, double *p, double *const vel[3], const mif_function *grad_p,
                      double dt);

/* The spectral solve alone on a packed owner-point array (i fastest), in
 * place (src/PressureEquation.cpp:65-264). */
void orc_poisson_packed(const orc_grid *g, double *b);

/* The spectral solve with cell-centred Neumann axes (no reference
 * counterpart): on an axis with cell[a] != 0 (a wall axis) the unknowns are
 * the N_g - 1 cell centres, the ghosts mirror the first/last cell and the
 * transform pair is the DCT-II / DCT-III.  b is the packed array (extents
 * N_g - 1 on cell-centred axes, npts elsewhere), solved in place with a zero
 * constant mode. */
void orc_poisson_packed_cell(const orc_grid *g, const int cell[3], double *b);
/* X_k = sum_j x_j cos(pi k (2j + 1) / (2n)) and its inverse up to the factor
 * n, y_j = X_0 + 2 sum_{k>=1} X_k cos(pi k (2j + 1) / (2n)); direct sums. */
void orc_dct2_direct(long n, const double *in, double *out);
void orc_dct3_direct(long n, const double *in, double *out);

/* The same solve with every 1-D transform done by O(N^2) direct sums: the
 * same mathematics in another FP64 operation order, to measure how much two
 * correct FP64 solves differ (not thread-safe against concurrent calls of
 * orc_poisson_packed). */
void orc_poisson_packed_direct(const orc_grid *g, double *b);
/* Every later solve (orc_solve_pressure, orc_timestep, ...) with direct-sum
 * transforms (on = 1) or the FFTs (on = 0, the default). */
void orc_set_direct_transforms(int on);

/* adjust_pressure (src/PressureEquation.cpp:288-343). */
void orc_adjust_pressure(const orc_grid *g, double *p, const mif_function *fn,
                         double t);

/* ErrorL1/L2/LInfNorm (src/Norms.cpp:11-120); out = {L1, L2, Linf}. */
void orc_error_norms_velocity(const orc_grid *g, double *const vel[3],
                              const mif_function *fn, double t, double out[3]);
void orc_error_norms_pressure(const orc_grid *g, const double *p,
                              const mif_function *fn, double t, double out[3]);

/* Pressure gradient field on interior points (test/full_test.cpp:141-142). */
void orc_pressure_gradient(const orc_grid *g, const double *p,
                           double *const out[3]);

/* Momentum right-hand side of one component at one point
 * (include/MomentumEquation.h:43-255); exposed for stencil tests. */
double orc_momentum_rhs(const orc_grid *g, double *const vel[3], int comp,
                        long i, long j, long k);

/* 1-D transforms with FFTW's definitions (FFTW_R2HC halfcomplex order,
 * FFTW_HC2R, FFTW_REDFT00), O(N log N) and O(N^2) direct forms. */
void orc_r2hc(long n, const double *in, double *out);
void orc_hc2r(long n, const double *in, double *out);
void orc_redft00(long n, const double *in, double *out);
void orc_r2hc_direct(long n, const double *in, double *out);
void orc_hc2r_direct(long n, const double *in, double *out);
void orc_redft00_direct(long n, const double *in, double *out);

/* Adjoint of one timestep() (no reference counterpart): given the three
 * stencil sources the forward step read (un = u^n, y2 = Y2 after its
 * projection, y3 = Y3 after its projection) and the adjoints of (u^{n+1},
 * p^{n+1}) in ubar, pbar, overwrite them with the adjoints of (u^n, p^n) and
 * add the adjoint of the body force to fbar.  Homogeneous pressure-increment
 * conditions (timestep, not timestep_nhn); Dirichlet data are constants of
 * the map and receive no adjoint. */
void orc_timestep_adjoint(const orc_grid *g, double *const un[3], double *const y2[3], double *const y3[3],
                          double *const ubar[3], double *pbar, double fbar[3]);

/* Number of OpenMP threads the oracle uses (1 when built without OpenMP). */
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif

#endif
