// Triples of host callbacks (VectorFunction / TimeVectorFunction,
// include/VectorFunction.h:15-53 of the reference), each optionally paired
// with a closed-form descriptor the device can evaluate.
//
// Device kernels cannot call a host std::function.  A TimeVectorFunction
// built from one of the reference's named free functions (u_exact, ...,
// forcing_x, ..., exact_u_t1, ...) is recognised and carries the matching
// mif_function descriptor; set_time and get_difference_over_time keep it.
// A descriptor can also be given directly.  Without one, the callbacks are
// evaluated on the host: boundary data (timestep, timestep_nhn,
// timestep_velocity, apply_bc, the Neumann data of the pressure solve) into
// O(N^2) face tables per stage, and set, adjust_pressure and the error norms
// into whole fields.
#ifndef MIF_SHIM_VECTOR_FUNCTION_H
#define MIF_SHIM_VECTOR_FUNCTION_H

#include <array>
#include <functional>

#include "Real.h"
#include "mif_gpu.h"

namespace mif {

using SpaceFunction = std::function<Real(Real, Real, Real)>;
using TimeFunction = std::function<Real(Real, Real, Real, Real)>;

class VectorFunction {
 public:
  VectorFunction(const SpaceFunction f_u, const SpaceFunction f_v, const SpaceFunction f_w);
  // The descriptor evaluated at a fixed time.
  VectorFunction(const mif_function &descriptor, Real time);
  VectorFunction(const VectorFunction &other);
  VectorFunction &operator=(const VectorFunction &) = delete;

  const SpaceFunction f_u;
  const SpaceFunction f_v;
  const SpaceFunction f_w;
  const std::array<const SpaceFunction *, 3> components;

  VectorFunction operator+(const VectorFunction &other) const;
  VectorFunction operator*(const Real scalar) const;

  bool has_descriptor() const { return has_desc_; }
  // The descriptor with its Reynolds number read now (the generated
  // functions read the global Reynolds at call time).
  mif_function descriptor() const;
  Real time() const { return time_; }

 private:
  bool has_desc_ = false;
  mif_function desc_{};
  Real time_ = 0;
};

class TimeVectorFunction {
 public:
  TimeVectorFunction(const TimeFunction f_u, const TimeFunction f_v, const TimeFunction f_w);
  explicit TimeVectorFunction(const mif_function &descriptor);
  TimeVectorFunction(const TimeVectorFunction &other);
  TimeVectorFunction &operator=(const TimeVectorFunction &) = delete;

  const TimeFunction f_u;
  const TimeFunction f_v;
  const TimeFunction f_w;
  const std::array<const TimeFunction *, 3> components;

  // The function at a fixed time (src/VectorFunction.cpp:47-51).
  VectorFunction set_time(Real time) const;
  // f(time_2) - f(time_1) (src/VectorFunction.cpp:53-61).
  VectorFunction get_difference_over_time(Real time_1, Real time_2) const;

  bool has_descriptor() const { return has_desc_; }
  mif_function descriptor() const;

 private:
  bool has_desc_ = false;
  mif_function desc_{};
};

// Descriptor of a scalar callback when it is one of the named exact
// pressures (p_exact, p_exact_p_test); false otherwise.
bool describe_scalar(const TimeFunction &f, mif_function *out);

// The Reynolds number the generated functions read (the caller's global
// `Reynolds`, include/Manufactured.h:4), or 1 when the program defines none.
double current_reynolds();

}  // namespace mif

#endif  // MIF_SHIM_VECTOR_FUNCTION_H
