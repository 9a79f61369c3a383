// Stand-alone Poisson test field (generators/manufsol_pressure.py:36-39,
// include/ManufacturedPressure.h of the reference).
#ifndef MIF_SHIM_MANUFACTURED_PRESSURE_H
#define MIF_SHIM_MANUFACTURED_PRESSURE_H

double u_exact_p_test(double t, double x, double y, double z);
double v_exact_p_test(double t, double x, double y, double z);
double w_exact_p_test(double t, double x, double y, double z);
double p_exact_p_test(double t, double x, double y, double z);
double dp_dx_exact_p_test(double t, double x, double y, double z);
double dp_dy_exact_p_test(double t, double x, double y, double z);
double dp_dz_exact_p_test(double t, double x, double y, double z);

#endif  // MIF_SHIM_MANUFACTURED_PRESSURE_H
