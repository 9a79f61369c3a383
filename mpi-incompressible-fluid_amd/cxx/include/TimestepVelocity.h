// Velocity-only RK3 step with the manufactured forcing
// (include/TimestepVelocity.h:15-16 of the reference), on the device.
#ifndef MIF_SHIM_TIMESTEP_VELOCITY_ONLY_H
#define MIF_SHIM_TIMESTEP_VELOCITY_ONLY_H

#include "VelocityTensor.h"

namespace mif {

// The forcing is the one the reference compiles in (forcing_x/y/z of
// include/ManufacturedVelocity.h); velocity and velocity_buffer exchange
// storage as in src/TimestepVelocity.cpp:89.
void timestep_velocity(VelocityTensor &velocity, VelocityTensor &velocity_buffer, VelocityTensor &rhs_buffer,
                       const TimeVectorFunction &exact_velocity, Real t_n);

}  // namespace mif

#endif  // MIF_SHIM_TIMESTEP_VELOCITY_ONLY_H
