/*
 * mif_types.h -- plain-C types shared by the HIP library's C ABI (mif_gpu.h)
 * and by anything that describes a problem to it.
 *
 * The grid descriptor carries exactly the arguments of the reference's
 * domain-constants constructor (include/Constants.h:86-90,
 * src/Constants.cpp:58-62) for a single rank (Py = Pz = 1): node counts per
 * axis including both end nodes (for a periodic axis the count includes the
 * duplicated end node, include/Constants.h:14-16), domain sizes and origins,
 * Reynolds number, final time and number of steps (dt = final_time /
 * num_time_steps, src/Constants.cpp:69) and the periodic flags.
 */
#ifndef MIF_TYPES_H
#define MIF_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes returned by every entry point (0 = success). */
typedef enum mif_status {
  MIF_OK = 0,
  MIF_ERR_INVALID_ARGUMENT = 1, /* bad grid, null pointer, wrong field kind  */
  MIF_ERR_HIP = 2,              /* a HIP runtime call failed                 */
  MIF_ERR_OUT_OF_MEMORY = 3,    /* device allocation failed                  */
  MIF_ERR_NON_FINITE = 4,       /* a NaN/Inf was detected in a reduction     */
  MIF_ERR_UNSUPPORTED = 5,      /* valid request the library does not handle */
  MIF_ERR_NO_DEVICE = 6         /* no HIP device is visible                  */
} mif_status;

/* Staggering of a field, as the reference's StaggeringDirection
 * (include/StaggeredTensor.h:13-15): u lives on x faces, v on y faces, w on z
 * faces, pressure on nodes. */
typedef enum mif_staggering {
  MIF_STAGGER_X = 0,
  MIF_STAGGER_Y = 1,
  MIF_STAGGER_Z = 2,
  MIF_STAGGER_NONE = 3
} mif_staggering;

/* One rank's domain constants (reference Constants at Py = Pz = 1). */
typedef struct mif_grid {
  uint64_t Nx_global, Ny_global, Nz_global; /* nodes incl. both ends        */
  double x_size, y_size, z_size;            /* domain lengths               */
  double min_x, min_y, min_z;               /* domain origin                */
  double Re;                                /* Reynolds number              */
  double final_time;                        /* dt = final_time / num_steps  */
  uint32_t num_time_steps;
  int32_t periodic[3];                      /* 1 = periodic along x, y, z   */
} mif_grid;

/* Analytic functions the reference passes as std::function callbacks
 * (include/VectorFunction.h:15-53).  Device code cannot call a host
 * std::function, so the known free functions are named here and evaluated on
 * the device from their closed forms; component 0/1/2 selects the x/y/z entry
 * of a vector function, scalar functions use component 0. */
typedef enum mif_fn_kind {
  MIF_FN_ZERO = 0,
  /* Ethier-Steinman exact Navier-Stokes solution, a = pi/4, d = pi/2
   * (generators/manufsol.py:31-72, u_exact/v_exact/w_exact, p_exact,
   * dp_dx_exact/dp_dy_exact/dp_dz_exact in include/Manufactured.h:7-13). */
  MIF_FN_ES_VELOCITY = 1,
  MIF_FN_ES_PRESSURE = 2,
  MIF_FN_ES_GRAD_P = 3,
  /* Velocity-only manufactured field and its forcing
   * (generators/manufsol_velocity.py:55-57, include/ManufacturedVelocity.h). */
  MIF_FN_VT_VELOCITY = 4,
  MIF_FN_VT_FORCING = 5,
  /* Stand-alone Poisson test field (generators/manufsol_pressure.py:36-39,
   * include/ManufacturedPressure.h). */
  MIF_FN_PT_VELOCITY = 6,
  MIF_FN_PT_PRESSURE = 7,
  MIF_FN_PT_GRAD_P = 8,
  /* Lid-driven cavity boundary data (include/TestCaseBoundaries.h:18-56). */
  MIF_FN_LID_1 = 9,
  MIF_FN_LID_2 = 10,
  /* Constant vector (c[0], c[1], c[2]); a uniform body force. */
  MIF_FN_CONSTANT = 11,
  /* f(t,x,y,z) = scale * g(t,x,y,z) difference of two times of a base kind:
   * value = base(t1) - base(t0); used for TimeVectorFunction::
   * get_difference_over_time (src/VectorFunction.cpp:53-61). */
  MIF_FN_TIME_DIFFERENCE = 12
} mif_fn_kind;

typedef struct mif_function {
  int32_t kind;     /* mif_fn_kind                                          */
  int32_t base;     /* base kind for MIF_FN_TIME_DIFFERENCE                 */
  double Re;        /* the generated functions read a global Reynolds
                       number (include/Manufactured.h:4)                    */
  double c[3];      /* MIF_FN_CONSTANT values                               */
  double t0, t1;    /* MIF_FN_TIME_DIFFERENCE: base(t1) - base(t0)          */
} mif_function;

#ifdef __cplusplus
}
#endif

#endif /* MIF_TYPES_H */
