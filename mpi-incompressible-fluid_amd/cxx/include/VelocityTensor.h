// Velocity components and the velocity triple (UTensor / VTensor / WTensor
// / VelocityTensor, include/VelocityTensor.h:11-99 of the reference).  u, v
// and w are offset by half a cell towards negative x, y and z respectively.
#ifndef MIF_SHIM_VELOCITY_TENSOR_H
#define MIF_SHIM_VELOCITY_TENSOR_H

#include <array>

#include "StaggeredTensor.h"
#include "VectorFunction.h"

namespace mif {

class UTensor : public StaggeredTensor {
 public:
  explicit UTensor(const Constants &constants) : StaggeredTensor(constants, StaggeringDirection::x) {}
  Real evaluate_function_at_index(int i, int j, int k, const SpaceFunction &f) const override {
    return f(constants.min_x_global + constants.dx * (constants.base_i + i) - constants.dx_over_2,
             constants.min_y_global + constants.dy * (constants.base_j + j),
             constants.min_z_global + constants.dz * (constants.base_k + k));
  }
  Real evaluate_function_at_index(Real time, int i, int j, int k, const TimeFunction &f) const override {
    return f(time, constants.min_x_global + constants.dx * (constants.base_i + i) - constants.dx_over_2,
             constants.min_y_global + constants.dy * (constants.base_j + j),
             constants.min_z_global + constants.dz * (constants.base_k + k));
  }
};

class VTensor : public StaggeredTensor {
 public:
  explicit VTensor(const Constants &constants) : StaggeredTensor(constants, StaggeringDirection::y) {}
  Real evaluate_function_at_index(int i, int j, int k, const SpaceFunction &f) const override {
    return f(constants.min_x_global + constants.dx * (constants.base_i + i),
             constants.min_y_global + constants.dy * (constants.base_j + j) - constants.dy_over_2,
             constants.min_z_global + constants.dz * (constants.base_k + k));
  }
  Real evaluate_function_at_index(Real time, int i, int j, int k, const TimeFunction &f) const override {
    return f(time, constants.min_x_global + constants.dx * (constants.base_i + i),
             constants.min_y_global + constants.dy * (constants.base_j + j) - constants.dy_over_2,
             constants.min_z_global + constants.dz * (constants.base_k + k));
  }
};

class WTensor : public StaggeredTensor {
 public:
  explicit WTensor(const Constants &constants) : StaggeredTensor(constants, StaggeringDirection::z) {}
  Real evaluate_function_at_index(int i, int j, int k, const SpaceFunction &f) const override {
    return f(constants.min_x_global + constants.dx * (constants.base_i + i),
             constants.min_y_global + constants.dy * (constants.base_j + j),
             constants.min_z_global + constants.dz * (constants.base_k + k) - constants.dz_over_2);
  }
  Real evaluate_function_at_index(Real time, int i, int j, int k, const TimeFunction &f) const override {
    return f(time, constants.min_x_global + constants.dx * (constants.base_i + i),
             constants.min_y_global + constants.dy * (constants.base_j + j),
             constants.min_z_global + constants.dz * (constants.base_k + k) - constants.dz_over_2);
  }
};

class VelocityTensor {
 public:
  UTensor u;
  VTensor v;
  WTensor w;
  std::array<StaggeredTensor *, 3> components;
  const Constants &constants;

  explicit VelocityTensor(const Constants &constants);
  VelocityTensor(const VelocityTensor &) = delete;
  VelocityTensor &operator=(const VelocityTensor &) = delete;

  // Constant-time exchange of the values (src/VelocityTensor.cpp:13-26).
  void swap_data(VelocityTensor &other);

  // Fill the components from f (src/VelocityTensor.cpp:28-34): on the device
  // when f carries a descriptor, else by host evaluation of the callbacks.
  void set(const VectorFunction &f, bool include_border);

  // Dirichlet data on the non-periodic faces and periodic copies
  // (src/VelocityTensor.cpp:36-233), on the device; f is evaluated there
  // when it carries a descriptor, else on the host into face tables.
  void apply_bc(const VectorFunction &exact_velocity);

  // The three device fields, with newer host values uploaded first.
  std::array<mif_field *, 3> device() const;
  void mark_device_written() const;
};

}  // namespace mif

#endif  // MIF_SHIM_VELOCITY_TENSOR_H
