// A staggered field held on the device (StaggeredTensor,
// include/StaggeredTensor.h:30-160 of the reference).
//
// The values live in a device field of the domain's context.  Element access
// through operator() goes to a host mirror that is synchronised lazily:
//   - a const access downloads the device values first if they are newer;
//   - a non-const access does the same and marks the host copy as newer;
//   - device() uploads a newer host copy before handing the field to a
//     device routine, and the routine's caller marks the device as newer.
// Extents and coordinates follow the reference at one rank
// (src/StaggeredTensor.cpp:5-9, include/StaggeredTensor.h:95-110).
#ifndef MIF_SHIM_STAGGERED_TENSOR_H
#define MIF_SHIM_STAGGERED_TENSOR_H

#include <array>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "Constants.h"

namespace mif {

enum StaggeringDirection { x, y, z, none };

class StaggeredTensor {
 public:
  StaggeredTensor(const Constants &constants, const StaggeringDirection &staggering);
  StaggeredTensor(const StaggeredTensor &) = delete;
  StaggeredTensor &operator=(const StaggeredTensor &) = delete;
  virtual ~StaggeredTensor();

  const Constants &constants;
  StaggeringDirection staggering;

  const std::array<size_t, 3> &sizes() const { return sizes_; }

  Real &operator()(size_t i, size_t j, size_t k) {
    pull();
    state_ = HOST_NEWER;
    return host_[i + sizes_[0] * (j + sizes_[1] * k)];
  }
  Real operator()(size_t i, size_t j, size_t k) const {
    pull();
    return host_[i + sizes_[0] * (j + sizes_[1] * k)];
  }
  // Host copy of the values (x fastest); the caller may write through it.
  void *raw_data() {
    pull();
    state_ = HOST_NEWER;
    return host_.data();
  }

  // Periodic copies on the periodic axes (src/StaggeredTensor.cpp:221-256),
  // on the device.
  void apply_periodic_bc();

  // Evaluate f at index (i, j, k): node coordinates, shifted by half a cell
  // on the staggered axis in the derived classes.
  inline Real evaluate_function_at_index_unstaggered(Real time, int i, int j, int k,
                                                     const std::function<Real(Real, Real, Real, Real)> &f) const {
    return f(time, constants.min_x_global + constants.dx * (constants.base_i + i),
             constants.min_y_global + constants.dy * (constants.base_j + j),
             constants.min_z_global + constants.dz * (constants.base_k + k));
  }
  inline Real evaluate_function_at_index_unstaggered(int i, int j, int k,
                                                     const std::function<Real(Real, Real, Real)> &f) const {
    return f(constants.min_x_global + constants.dx * (constants.base_i + i),
             constants.min_y_global + constants.dy * (constants.base_j + j),
             constants.min_z_global + constants.dz * (constants.base_k + k));
  }
  virtual Real evaluate_function_at_index(Real time, int i, int j, int k,
                                          const std::function<Real(Real, Real, Real, Real)> &f) const {
    return evaluate_function_at_index_unstaggered(time, i, j, k, f);
  }
  virtual Real evaluate_function_at_index(int i, int j, int k, const std::function<Real(Real, Real, Real)> &f) const {
    return evaluate_function_at_index_unstaggered(i, j, k, f);
  }

  // Fill every point (include_border) or the interior points with f
  // (src/StaggeredTensor.cpp:205-219); f is a host callback, so the values
  // are computed on the host and uploaded on the next device use.
  void set(const std::function<Real(Real, Real, Real)> &f, bool include_border);

  // Exchange the values with another tensor of the same staggering in
  // constant time (device pointers and host mirrors).
  void swap_data(StaggeredTensor &other);

  // Text dump and load (Tensor::dump / Tensor::load, include/Tensor.h:433-476):
  // dump writes "Tensor (row major). Dimensions: [ sx sy sz ]" and the values
  // in storage order (x fastest) between brackets, with the stream's default
  // formatting (6 significant digits, as the reference prints them); load
  // reads whitespace-separated values in storage order, either a plain list
  // or a dump's body (the header is skipped), and leaves the remaining values
  // unchanged when the file holds fewer.  Both return 1, or print "Error
  // opening <file> file." and return 0.
  uint8_t dump(const std::string &file_name) const;
  uint8_t load(const std::string &file_name);
  // Exact restart files: a 48-byte header ("MIFTNSR1", the three extents and
  // the staggering as 64-bit integers, the value size) followed by the
  // values as little-endian doubles in storage order.  load_binary checks
  // the extents and the staggering; both return 1 on success, 0 otherwise.
  uint8_t dump_binary(const std::string &file_name) const;
  uint8_t load_binary(const std::string &file_name);

  // The device field, with any newer host values uploaded first.
  mif_field *device() const;
  // Record that a device routine has written the field.
  void mark_device_written() const { state_ = DEVICE_NEWER; }

 private:
  enum State { SYNCED, HOST_NEWER, DEVICE_NEWER };
  void pull() const;

  std::array<size_t, 3> sizes_;
  mif_field *field_ = nullptr;
  mutable std::vector<Real> host_;
  mutable State state_ = SYNCED;
};

}  // namespace mif

#endif  // MIF_SHIM_STAGGERED_TENSOR_H
