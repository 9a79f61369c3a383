// Velocity-only manufactured field and its forcing
// (generators/manufsol_velocity.py:32-57, include/ManufacturedVelocity.h of
// the reference).
#ifndef MIF_SHIM_MANUFACTURED_VELOCITY_H
#define MIF_SHIM_MANUFACTURED_VELOCITY_H

extern double Reynolds;

double u_exact_v_test(double t, double x, double y, double z);
double v_exact_v_test(double t, double x, double y, double z);
double w_exact_v_test(double t, double x, double y, double z);
double forcing_x(double t, double x, double y, double z);
double forcing_y(double t, double x, double y, double z);
double forcing_z(double t, double x, double y, double z);

#endif  // MIF_SHIM_MANUFACTURED_VELOCITY_H
