// Error norms against an exact solution (include/Norms.h:19-45 of the
// reference), reduced on the device.  A named exact solution is evaluated on
// the device; any other callback is tabulated on the host first.
#ifndef MIF_SHIM_NORMS_H
#define MIF_SHIM_NORMS_H

#include "VectorFunction.h"
#include "VelocityTensor.h"

namespace mif {

Real ErrorL1Norm(const VelocityTensor &velocity, const TimeVectorFunction &exact_velocity, Real time);
Real ErrorL2Norm(const VelocityTensor &velocity, const TimeVectorFunction &exact_velocity, Real time);
Real ErrorLInfNorm(const VelocityTensor &velocity, const TimeVectorFunction &exact_velocity, Real time);

Real ErrorL1Norm(const StaggeredTensor &pressure, const std::function<Real(Real, Real, Real, Real)> &exact_pressure,
                 Real time);
Real ErrorL2Norm(const StaggeredTensor &pressure, const std::function<Real(Real, Real, Real, Real)> &exact_pressure,
                 Real time);
Real ErrorLInfNorm(const StaggeredTensor &pressure, const std::function<Real(Real, Real, Real, Real)> &exact_pressure,
                   Real time);

// One rank: the local error is the global error.
Real accumulate_error_mpi_l1(Real local_error, const Constants &constants);
Real accumulate_error_mpi_l2(Real local_error, const Constants &constants);
Real accumulate_error_mpi_linf(Real local_error, const Constants &constants);

}  // namespace mif

#endif  // MIF_SHIM_NORMS_H
