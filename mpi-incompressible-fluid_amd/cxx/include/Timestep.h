// One full RK3 + projection step (include/Timestep.h:16-27 of the
// reference), on the device.
#ifndef MIF_SHIM_TIMESTEP_H
#define MIF_SHIM_TIMESTEP_H

#include "PressureTensor.h"
#include "VelocityTensor.h"

namespace mif {

// Homogeneous Neumann / periodic pressure increment (src/Timestep.cpp:
// 97-141).  exact_velocity gives the Dirichlet data: evaluated on the
// device when it carries a descriptor, else on the host at the three stage
// times into face tables (O(N^2) callback evaluations per stage).
void timestep(VelocityTensor &velocity, VelocityTensor &velocity_buffer, VelocityTensor &velocity_buffer_2,
              const TimeVectorFunction &exact_velocity, Real t_n, StaggeredTensor &pressure,
              StaggeredTensor &pressure_buffer, PressureTensor &pressure_solver_buffer);

// Non-homogeneous Neumann pressure increment from the exact pressure
// gradient (src/Timestep.cpp:143-156).
void timestep_nhn(VelocityTensor &velocity, VelocityTensor &velocity_buffer, VelocityTensor &velocity_buffer_2,
                  const TimeVectorFunction &exact_velocity, const TimeVectorFunction &exact_pressure_gradient,
                  Real t_n, StaggeredTensor &pressure, StaggeredTensor &pressure_buffer,
                  PressureTensor &pressure_solver_buffer);

}  // namespace mif

#endif  // MIF_SHIM_TIMESTEP_H
