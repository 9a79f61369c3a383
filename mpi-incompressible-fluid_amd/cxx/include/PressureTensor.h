// The pressure solver's packed buffer (PressureTensor, include/
// PressureTensor.h of the reference): the owner points of a pressure field,
// x fastest.  The device solve packs its own copy; this host array serves
// callers that move data between the two layouts themselves.
#ifndef MIF_SHIM_PRESSURE_TENSOR_H
#define MIF_SHIM_PRESSURE_TENSOR_H

#include "PressureSolverStructures.h"
#include "StaggeredTensor.h"
#include "Tensor.h"

namespace mif {

class PressureTensor : public Tensor<Real, 1U, int> {
 public:
  explicit PressureTensor(PressureSolverStructures &structures);
  PressureTensor(const PressureTensor &) = delete;

  PressureSolverStructures &structures;
  const int max_size;

  // Owner points of other into this buffer (src/PressureTensor.cpp:16-19).
  void copy_from_staggered(const StaggeredTensor &other);
  // The reverse, then the periodic copies (src/PressureTensor.cpp:21-34);
  // base_tag is accepted for the reference's signature (one rank: no
  // messages).
  void copy_to_staggered(StaggeredTensor &other, int base_tag) const;
};

}  // namespace mif

#endif  // MIF_SHIM_PRESSURE_TENSOR_H
