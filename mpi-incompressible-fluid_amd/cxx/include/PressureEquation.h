// Pressure Poisson solves and the pressure constant (include/
// PressureEquation.h:10-26 of the reference), on the device.
#ifndef MIF_SHIM_PRESSURE_EQUATION_H
#define MIF_SHIM_PRESSURE_EQUATION_H

#include "PressureTensor.h"
#include "VelocityTensor.h"

namespace mif {

void solve_pressure_equation_homogeneous_periodic(StaggeredTensor &pressure, PressureTensor &pressure_buffer,
                                                  const VelocityTensor &velocity, Real dt);

// exact_pressure_gradient is evaluated on the boundary planes on the device
// when it carries a descriptor, else on the host into a Neumann face table.
void solve_pressure_equation_non_homogeneous_neumann(StaggeredTensor &pressure, PressureTensor &pressure_buffer,
                                                     const VelocityTensor &velocity,
                                                     const VectorFunction &exact_pressure_gradient, Real dt);

// Shift the pressure by the mean of (exact - pressure) over the owner points
// (src/PressureEquation.cpp:288-343).  The callback is tabulated on the host
// and the reduction and shift run on the device.
void adjust_pressure(StaggeredTensor &pressure, const std::function<Real(Real, Real, Real)> &exact_pressure);

}  // namespace mif

#endif  // MIF_SHIM_PRESSURE_EQUATION_H
