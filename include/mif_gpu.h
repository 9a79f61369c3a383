/*
 * mif_gpu.h -- C ABI of the MI355X (gfx950) incompressible Navier-Stokes
 * timestep library (libmif_gpu.so).
 *
 * The entry points replace the reference solver's single-rank hot path: its
 * tensors become device fields (mif_field), its domain constants plus
 * pressure-solver structures become a context (mif_ctx), and each solver
 * function keeps the reference's arguments and meaning.  Plain pointers and
 * sizes only; no C++ types cross the boundary.  Every function returns a
 * mif_status (0 = success) and records a message retrievable with
 * mif_last_error().
 *
 * Host arrays passed to upload/download use the reference tensor layout:
 * extents (sx, sy, sz) of the staggering, index i + sx*(j + sy*k), FP64
 * (include/Tensor.h:127-141,219-238).  Device storage is private (padded
 * pitch; FP32 in the mixed-precision form).
 *
 * Solver calls are enqueued on the context's HIP stream and return without
 * waiting; uploads, downloads, norms and mif_ctx_synchronize() wait.  A
 * context is not thread-safe; separate contexts may be used from separate
 * threads.
 */
#ifndef MIF_GPU_H
#define MIF_GPU_H

#include "mif_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mif_ctx mif_ctx;
typedef struct mif_field mif_field;

/* Storage/arithmetic precision of a context. */
typedef enum mif_precision {
  MIF_PRECISION_FP64 = 0,  /* plain form: FP64 fields, stencils, transforms  */
  MIF_PRECISION_MIXED = 1  /* FP32 fields/stencils/transforms, FP64 defect
                              correction of the pressure solve               */
} mif_precision;

/* ---- context ------------------------------------------------------------
 * Replaces Constants::Constants (src/Constants.cpp:58-120) at Py = Pz = 1
 * together with PressureSolverStructures::PressureSolverStructures
 * (src/PressureSolverStructures.cpp:13-69): transform plans, eigenvalue
 * tables and scratch live in the context.  precision is a mif_precision;
 * device is the HIP device ordinal (-1 = current). */
int mif_ctx_create(const mif_grid *grid, int precision, int device, mif_ctx **out);
/* A context for stand-alone Poisson solves with cell-centred Neumann axes
 * (no counterpart in the reference, whose Neumann axes are node-centred:
 * FFTW_REDFT00, src/PressureSolverStructures.cpp:23-28).  On an axis with
 * cell_centred[a] != 0 (walls, not periodic) the unknowns are the N_g - 1
 * cell centres between the N_g nodes of the grid, the walls reflect the
 * first/last cell (phi_{-1} = phi_0), the operator's rows are (-1, 1),
 * (1, -2, 1), (1, -1) over h^2, and the transform pair is the DCT-II /
 * DCT-III with eigenvalues 2 (cos(pi k / N) - 1) / h^2.  The y axis only,
 * with 64, 128, 256, 512 or 1024 cells (else MIF_ERR_UNSUPPORTED).  Such a
 * context has no staggered fields (mif_field_create returns
 * MIF_ERR_UNSUPPORTED); it serves mif_poisson_set_rhs / mif_poisson_run /
 * mif_poisson_get_solution, whose packed arrays then hold N_g - 1 points
 * along that axis. */
int mif_ctx_create_cell_centred(const mif_grid *grid, const int cell_centred[3], int precision, int device,
                                mif_ctx **out);
/* Frees everything the context owns (PressureSolverStructures::
 * ~PressureSolverStructures, src/PressureSolverStructures.cpp:72-97).  Fields
 * must be destroyed first. */
int mif_ctx_destroy(mif_ctx *ctx);
/* Waits for all work enqueued on the context's stream. */
int mif_ctx_synchronize(mif_ctx *ctx);
/* Message of the last failure on this context (or of the last failed
 * mif_ctx_create when ctx is NULL). */
const char *mif_last_error(const mif_ctx *ctx);
/* Derived constants: dt (src/Constants.cpp:69) and the node extents
 * Nx, Ny, Nz including periodic ghosts (src/Constants.cpp:80-82). */
int mif_ctx_info(const mif_ctx *ctx, double *dt, uint64_t node_extents[3]);
/* Number of FP64 defect-correction sweeps of the mixed form's pressure solve
 * (default 1; 0..4).  Ignored in the FP64 form. */
int mif_ctx_set_defect_sweeps(mif_ctx *ctx, int sweeps);
/* Defect-correction sweep count of the context. */
int mif_ctx_get_defect_sweeps(const mif_ctx *ctx, int *sweeps);
/* Which pressure solves of mif_timestep and mif_timestep_nhn (and their
 * _tabulated forms) take the defect sweeps (no reference counterpart; the reference solves in one
 * precision).  MIF_DEFECT_EVERY_STAGE (the default): all three.
 * MIF_DEFECT_LAST_STAGE: the solve of stage 3 only, whose increment
 * completes p^{n+1} and whose projection leaves u^{n+1}; the solves of
 * stages 1 and 2 use the FP32 transforms alone.  Stand-alone solves
 * (mif_solve_pressure_*, mif_poisson_solve, mif_poisson_run) always take
 * the sweeps. */
typedef enum mif_defect_schedule {
  MIF_DEFECT_EVERY_STAGE = 0,
  MIF_DEFECT_LAST_STAGE = 1
} mif_defect_schedule;
int mif_ctx_set_defect_schedule(mif_ctx *ctx, int schedule);

/* Transforms of the mixed form's pressure solve (no reference counterpart;
 * the reference solves in one precision).  FP32 transforms amplify their
 * rounding by the condition number kappa of the discrete Laplacian, and
 * both the incremental projection and the defect correction stop
 * contracting when that amplification nears one; MIF_SOLVE_AUTO keeps FP32
 * transforms up to kappa = 6e5 (the 512^3 channel has 4.0e5) and uses FP64
 * transforms (FP32 fields and stencils still) above it (768^3: 8.9e5). */
typedef enum mif_solve_precision {
  MIF_SOLVE_AUTO = 0,
  MIF_SOLVE_FP32 = 1,
  MIF_SOLVE_FP64 = 2
} mif_solve_precision;
int mif_ctx_set_solve_precision(mif_ctx *ctx, int mode);
/* The transforms the context will use (MIF_SOLVE_FP32 or MIF_SOLVE_FP64) and,
 * if kappa is not NULL, the condition-number estimate max|lambda| /
 * min|lambda != 0| of the discrete Laplacian. */
int mif_ctx_get_solve_precision(const mif_ctx *ctx, int *effective, double *kappa);

/* ---- fields ---------------------------------------------------------------
 * Replaces StaggeredTensor(const Constants&, StaggeringDirection)
 * (include/StaggeredTensor.h:32, src/StaggeredTensor.cpp:5-9); staggering is
 * a mif_staggering.  The field is zero-initialised. */
int mif_field_create(mif_ctx *ctx, int staggering, mif_field **out);
int mif_field_destroy(mif_field *f);
/* Tensor::sizes() (include/Tensor.h:151-153). */
int mif_field_extents(const mif_field *f, uint64_t extents[3]);
/* Copy a host array in the reference layout to the device (Tensor raw_data
 * writes, include/Tensor.h:115). */
int mif_field_upload(mif_field *f, const double *host);
/* Copy the field to a host array in the reference layout. */
int mif_field_download(const mif_field *f, double *host);
/* The same two copies with a device array in the reference layout (FP64),
 * e.g. a torch tensor's storage; synchronise the context's stream. */
int mif_field_upload_device(mif_field *f, const double *device_src);
int mif_field_download_device(const mif_field *f, double *device_dst);
/* Tensor::swap_data (include/Tensor.h:107-109): exchange device storage. */
int mif_field_swap(mif_field *a, mif_field *b);
/* StaggeredTensor::set (src/StaggeredTensor.cpp:205-219) with a known
 * function at time t (component comp of a vector function). */
int mif_field_set(mif_field *f, const mif_function *fn, int comp, double t, int include_border);
/* StaggeredTensor::apply_periodic_bc (src/StaggeredTensor.cpp:221-256). */
int mif_field_apply_periodic_bc(mif_field *f);

/* ---- velocity -------------------------------------------------------------
 * VelocityTensor::apply_bc (src/VelocityTensor.cpp:36-233) with the exact
 * velocity at time t (TimeVectorFunction::set_time, src/VectorFunction.cpp:
 * 47-51).  vel = {u, v, w}. */
int mif_velocity_apply_bc(mif_field *const vel[3], const mif_function *exact_velocity, double t);

/* ---- boundary data from host callbacks ---------------------------------------
 * The reference evaluates its boundary data through std::function callbacks
 * (include/VectorFunction.h:15-53) at every stage.  The entry points above
 * take closed-form descriptors the device evaluates; a callback without one
 * is evaluated on the host into face tables that these variants take:
 *   velocity face table: for u, v, w in turn, the faces z-, z+, y-, y+, x-, x+
 *     of the non-periodic axes, each a plane over the component's extents
 *     along the two other axes, lower axis fastest (z faces i + ex*j, y faces
 *     i + ex*k, x faces j + ey*k), holding the value that
 *     VelocityTensor::apply_bc writes at that point of that face
 *     (src/VelocityTensor.cpp:36-233): the exact value for a tangential
 *     component, the extrapolated value for the normal one;
 *   Neumann table: the six faces of the node extents in the same order and
 *     indexing, holding the normal component of the pressure-gradient data at
 *     each wall node (x component on x faces, ...), as
 *     apply_non_homogeneous_neumann reads it (src/PressureEquation.cpp:10-57).
 * mif_bc_table_size gives the element count of one table of each kind.  The
 * tables are copied before the call returns.  Edge points of a face that a
 * later face overwrites, and faces of periodic axes, follow the reference's
 * order as in the descriptor forms. */
int mif_bc_table_size(const mif_ctx *ctx, int64_t *velocity_faces, int64_t *neumann_faces);
/* VelocityTensor::apply_bc with a velocity face table. */
int mif_velocity_apply_bc_tabulated(mif_field *const vel[3], const double *faces);

/* ---- timesteps --------------------------------------------------------------
 * timestep (include/Timestep.h:16-19, src/Timestep.cpp:151-156): one RK3 step
 * with an incremental projection per stage, homogeneous Neumann/periodic
 * pressure increment.  vel, buf, buf2 are {u, v, w} triples. */
int mif_timestep(mif_field *const vel[3], mif_field *const buf[3], mif_field *const buf2[3],
                 const mif_function *exact_velocity, double t_n, mif_field *pressure,
                 mif_field *pressure_buffer);
/* nsteps successive mif_timestep calls at t_n, t_n + dt, ...: the result is
 * that of the loop (to rounding).  No reference counterpart (the reference's
 * driver loops over timestep, src/main.cpp); a step's last projection is
 * applied by the next step's first stage kernel instead of a pass of its
 * own, except after the last step.  With stage records on
 * (mif_ctx_set_stage_records) it is that loop. */
int mif_timestep_n(mif_field *const vel[3], mif_field *const buf[3], mif_field *const buf2[3],
                   const mif_function *exact_velocity, double t_n, mif_field *pressure,
                   mif_field *pressure_buffer, int nsteps);
/* timestep_nhn (include/Timestep.h:23-27, src/Timestep.cpp:143-149): the
 * same with non-homogeneous Neumann data from the exact pressure gradient
 * (all-wall domains only). */
int mif_timestep_nhn(mif_field *const vel[3], mif_field *const buf[3], mif_field *const buf2[3],
                     const mif_function *exact_velocity, const mif_function *exact_pressure_gradient,
                     double t_n, mif_field *pressure, mif_field *pressure_buffer);
/* timestep_velocity (include/TimestepVelocity.h:15-16,
 * src/TimestepVelocity.cpp:60-90): velocity-only RK3 with forcing; on return
 * vel and buf have swapped storage as in the reference. */
int mif_timestep_velocity(mif_field *const vel[3], mif_field *const buf[3], mif_field *const rhs[3],
                          const mif_function *exact_velocity, const mif_function *forcing, double t_n);
/* The three timesteps with boundary data from host callbacks: `faces` holds
 * three velocity face tables back to back, the exact velocity at the stage
 * times (timestep: t_n + 64/120 dt, then + 16/120 dt, then t_n + dt,
 * src/Timestep.cpp:99-104; timestep_velocity: t_n + 8/15 dt, t_n + 2/3 dt,
 * t_n + dt, src/TimestepVelocity.cpp:60-90); `neumann` holds three Neumann
 * tables of the data each stage's solve adds, exact_pressure_gradient.
 * get_difference_over_time(new_time, prev_time) = grad p(prev_time) -
 * grad p(new_time) (src/Timestep.cpp:89-95, src/VectorFunction.cpp:53-61). */
int mif_timestep_tabulated(mif_field *const vel[3], mif_field *const buf[3], mif_field *const buf2[3],
                           const double *faces, double t_n, mif_field *pressure, mif_field *pressure_buffer);
int mif_timestep_nhn_tabulated(mif_field *const vel[3], mif_field *const buf[3], mif_field *const buf2[3],
                               const double *faces, const double *neumann, double t_n, mif_field *pressure,
                               mif_field *pressure_buffer);
int mif_timestep_velocity_tabulated(mif_field *const vel[3], mif_field *const buf[3], mif_field *const rhs[3],
                                    const double *faces, const mif_function *forcing, double t_n);

/* ---- pressure -------------------------------------------------------------
 * solve_pressure_equation_homogeneous_periodic (include/PressureEquation.h:
 * 10-13, src/PressureEquation.cpp:266-274): pressure <- solution of
 * L phi = div(vel)/dt, with periodic ghosts. */
int mif_solve_pressure_homogeneous_periodic(mif_field *pressure, mif_field *const vel[3], double dt);
/* solve_pressure_equation_non_homogeneous_neumann (include/PressureEquation.h:
 * 16-21, src/PressureEquation.cpp:276-286); exact_pressure_gradient is
 * evaluated at time t (VectorFunction after set_time). */
int mif_solve_pressure_non_homogeneous_neumann(mif_field *pressure, mif_field *const vel[3],
                                               const mif_function *exact_pressure_gradient, double t,
                                               double dt);
/* The same with the Neumann data as one host table (see above). */
int mif_solve_pressure_non_homogeneous_neumann_tabulated(mif_field *pressure, mif_field *const vel[3],
                                                         const double *neumann, double dt);
/* The spectral solve alone (src/PressureEquation.cpp:65-264 between
 * copy_from_staggered and copy_to_staggered): the right-hand side is read
 * from the owner points of `pressure`, the solution replaces it (owner
 * points and periodic ghosts). */
int mif_poisson_solve(mif_field *pressure);
/* The spectral solve on a caller-given right-hand side (the owner-point
 * array the reference packs with PressureTensor::copy_from_staggered,
 * src/PressureTensor.cpp:16-19: npts_x * npts_y * npts_z FP64 values, x
 * fastest).  set_rhs uploads and keeps it; run solves from it, device
 * resident (FP64 form: the reference's solve, src/PressureEquation.cpp:65-264;
 * mixed form: FP32 transforms plus the context's FP64 defect sweeps, with the
 * right-hand side first projected onto the range of the Laplacian);
 * get_solution downloads phi in the same layout. */
int mif_poisson_set_rhs(mif_ctx *ctx, const double *b);
int mif_poisson_run(mif_ctx *ctx);
int mif_poisson_get_solution(mif_ctx *ctx, double *phi);
/* adjust_pressure (include/PressureEquation.h:25-26,
 * src/PressureEquation.cpp:288-343) with the exact pressure at time t. */
int mif_adjust_pressure(mif_field *pressure, const mif_function *exact_pressure, double t);
/* The same with the exact pressure given as a table: `exact` is a node field
 * (MIF_STAGGER_NONE) of the same context holding p_exact at every point.  For
 * callers whose exact solution is a host callback (the reference passes a
 * std::function, e.g. a lambda fixing the time, test/full_test.cpp:146). */
int mif_adjust_pressure_tabulated(mif_field *pressure, const mif_field *exact);

/* ---- diagnostics ------------------------------------------------------------
 * ErrorL1Norm/ErrorL2Norm/ErrorLInfNorm of the velocity (include/Norms.h:
 * 19-29, src/Norms.cpp:11-85); out = {L1, L2, Linf}. */
int mif_error_norms_velocity(mif_field *const vel[3], const mif_function *exact_velocity, double t,
                             double out[3]);
/* Velocity norms with the exact components given as three node fields holding
 * the exact u, v, w at the nodes (the points src/Norms.cpp:24-44 evaluates
 * them at); for host-callback exact solutions. */
int mif_error_norms_velocity_tabulated(mif_field *const vel[3], mif_field *const exact_at_nodes[3], double out[3]);
/* The scalar versions (include/Norms.h:32-34, src/Norms.cpp:88-120). */
int mif_error_norms_pressure(const mif_field *pressure, const mif_function *exact_pressure, double t,
                             double out[3]);
/* Scalar norms against a node-field table of the exact pressure. */
int mif_error_norms_pressure_tabulated(const mif_field *pressure, const mif_field *exact, double out[3]);
/* Pressure-gradient field on interior points (pressure_gradient_{u,v,w},
 * include/PressureGradient.h:9-25, as filled in test/full_test.cpp:141-142). */
int mif_pressure_gradient(const mif_field *pressure, mif_field *const out[3]);

/* ---- adjoint (no reference counterpart) -----------------------------------
 * Constant body force f added to the momentum right-hand side N of every
 * stage of mif_timestep (default 0, which is the reference's equation). */
int mif_ctx_set_body_force(mif_ctx *ctx, const double force[3]);
/* keep != 0: mif_timestep (and mif_timestep_nhn) leave Y2 and Y3 after their
 * projections in the `buf` and `buf2` triples, the stage records
 * mif_timestep_adjoint reads.  By default (keep = 0) the projections of
 * stages 1 and 2 are fused into the next stage's kernel, which reads the
 * buffers before their projection, and the buffers then hold Y2 and Y3
 * without it; u^{n+1} and p^{n+1} are the same either way (to rounding).
 * No counterpart in the reference, whose timestep leaves its buffers
 * unspecified to callers. */
int mif_ctx_set_stage_records(mif_ctx *ctx, int keep);
/* Adjoint of one mif_timestep (homogeneous pressure-increment conditions).
 * un, y2, y3 are the velocity fields the forward step used as stencil
 * sources: u^n, and the `buf` and `buf2` triples as mif_timestep leaves them
 * with stage records on (Y2 and Y3 after their projections, see
 * mif_ctx_set_stage_records).  On entry ubar and pbar hold the
 * adjoints of (u^{n+1}, p^{n+1}); on return those of (u^n, p^n).
 * force_bar[3] is incremented by the adjoint of the body force, which
 * synchronises the stream; with force_bar == NULL it is added to an
 * accumulator in the context instead and the call does not synchronise.
 * Dirichlet data are constants of the map. */
int mif_timestep_adjoint(mif_field *const un[3], mif_field *const y2[3], mif_field *const y3[3],
                         mif_field *const ubar[3], mif_field *pbar, double force_bar[3]);
/* Adds the context's force-adjoint accumulator (the calls above with
 * force_bar == NULL since the last take) to force_bar[3] and resets it.
 * Synchronises the stream. */
int mif_adjoint_force_take(mif_ctx *ctx, double force_bar[3]);
/* Device copy between fields of one staggering (Tensor copy). */
int mif_field_copy(mif_field *dst, const mif_field *src);
/* Sum over all points of the extents of a * b (FP64 accumulation). */
int mif_field_dot(const mif_field *a, const mif_field *b, double *out);
/* Values of one field at n points of its extents, given as (i, j, k)
 * triples in ijk[3n], written as FP64 into out[n].  No counterpart in the
 * reference, which reads its host tensors directly: the output writers
 * (writeVTK/writeDat, src/VTKDatExport.cpp) fetch only the points they
 * print through this.  Allocates device scratch per call; synchronises. */
int mif_field_gather(const mif_field *f, const int32_t *ijk, int64_t n, double *out);
/* The same gather with no device allocation: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving or output loops, harnesses that forbid hipMalloc).
 * The indices and values pass through a fixed device array of the library
 * (one per GPU, held until the process ends) instead of per-call buffers, so
 * n is limited to 1048576 points; larger n returns MIF_ERR_INVALID_ARGUMENT.
 * All calls on one GPU share that array: calls that overlap in time on one
 * GPU (contexts used from several threads) race.  Synchronises. */
int mif_field_gather_noalloc(const mif_field *f, const int32_t *ijk, int64_t n, double *out);

/* ---- measurement ------------------------------------------------------------
 * Kernel timing with HIP events recorded on the context's stream around
 * every launch of the kernels of one class (no reference counterpart; used
 * to report the roofline of a kernel from inside a run). */
typedef enum mif_kernel_class {
  MIF_KC_STAGE = 0,     /* fused RK3 stage (three components)              */
  MIF_KC_BC = 1,        /* boundary shell                                   */
  MIF_KC_DIV = 2,       /* divergence -> packed right-hand side            */
  MIF_KC_FFT_X = 3,     /* x forward/inverse transform passes              */
  MIF_KC_FFT_Y = 4,     /* y forward/inverse transform passes              */
  MIF_KC_FFT_Z = 5,     /* fused z forward/scale/inverse pass              */
  MIF_KC_UNPACK = 6,    /* packed solution -> pressure increment field     */
  MIF_KC_PROJECT = 7,   /* p += phi, u -= dt grad phi                      */
  MIF_KC_DEFECT = 8,    /* mixed form: FP64 residual / correction          */
  MIF_KC_FFT_X_DIV = 9, /* divergence fused into the x forward pass        */
  MIF_KC_FFT_X_UNPACK = 10, /* x inverse pass writing the pressure field   */
  MIF_KC_ADJOINT = 11,  /* whole adjoint step (all its kernels)            */
  MIF_KC_PRESSURE_RIM = 12, /* p += phi on the points a stage kernel with a fused
                               projection leaves out                        */
  MIF_KC_COUNT = 13
} mif_kernel_class;
int mif_ctx_set_timing(mif_ctx *ctx, int enable);
/* Total milliseconds and launch count of a class since timing was enabled
 * (synchronises the stream). */
int mif_ctx_kernel_time(mif_ctx *ctx, int kernel_class, double *total_ms, uint64_t *launches);
const char *mif_kernel_class_name(int kernel_class);

/* Measurement aid, no counterpart in the reference: the streaming ceiling of
 * a read/write mix on this device.  Allocates nread + nwrite device arrays of
 * `bytes` bytes each (0..8 of each, freed before return), times `reps`
 * launches of a kernel that reads every byte of the first nread arrays and
 * writes every byte of the other nwrite (16-byte accesses, grid-stride, on
 * the context's stream) at 4, 8, 16 and 32 workgroups per CU, and returns
 * the fastest average launch in microseconds.  Synchronises the stream. */
int mif_stream_probe(mif_ctx *ctx, int nread, int nwrite, int64_t bytes, int reps, double *us_per_launch);

#ifdef __cplusplus
}
#endif

#endif /* MIF_GPU_H */
