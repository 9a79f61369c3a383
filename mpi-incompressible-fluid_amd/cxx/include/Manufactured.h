// Ethier-Steinman exact solution of the Navier-Stokes equations, a = pi/4,
// d = pi/2 (the functions the reference generates from
// generators/manufsol.py:31-72 and declares in include/Manufactured.h).
// They read the caller's global Reynolds number.
#ifndef MIF_SHIM_MANUFACTURED_H
#define MIF_SHIM_MANUFACTURED_H

extern double Reynolds;

double u_exact(double t, double x, double y, double z);
double v_exact(double t, double x, double y, double z);
double w_exact(double t, double x, double y, double z);
double p_exact(double t, double x, double y, double z);
double dp_dx_exact(double t, double x, double y, double z);
double dp_dy_exact(double t, double x, double y, double z);
double dp_dz_exact(double t, double x, double y, double z);

#endif  // MIF_SHIM_MANUFACTURED_H
