// Backward differences of the pressure at the velocity points
// (include/PressureGradient.h:9-25 of the reference).  The point functions
// read the host mirror; pressure_gradient fills a whole velocity tensor on
// the device.
#ifndef MIF_SHIM_PRESSURE_GRADIENT_H
#define MIF_SHIM_PRESSURE_GRADIENT_H

#include "StaggeredTensor.h"
#include "VelocityTensor.h"

namespace mif {

inline Real pressure_gradient_u(const StaggeredTensor &p, const size_t i, const size_t j, const size_t k) {
  return (p(i, j, k) - p(i - 1, j, k)) * p.constants.one_over_dx;
}
inline Real pressure_gradient_v(const StaggeredTensor &p, const size_t i, const size_t j, const size_t k) {
  return (p(i, j, k) - p(i, j - 1, k)) * p.constants.one_over_dy;
}
inline Real pressure_gradient_w(const StaggeredTensor &p, const size_t i, const size_t j, const size_t k) {
  return (p(i, j, k) - p(i, j, k - 1)) * p.constants.one_over_dz;
}

// out = grad p on the interior points of each component (the loop of
// test/full_test.cpp:141-142), on the device.
void pressure_gradient(const StaggeredTensor &pressure, VelocityTensor &out);

}  // namespace mif

#endif  // MIF_SHIM_PRESSURE_GRADIENT_H
