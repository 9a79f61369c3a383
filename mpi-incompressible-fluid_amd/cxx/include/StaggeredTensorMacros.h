// Loop helpers over the points of a staggered tensor with the names and
// semantics of the reference's include/StaggeredTensorMacros.h: CODE sees
// the indices i, j, k.  include_border must be a constant expression.
#ifndef MIF_SHIM_STAGGERED_TENSOR_MACROS_H
#define MIF_SHIM_STAGGERED_TENSOR_MACROS_H

#include <array>
#include <cstddef>

// Every point (include_border) or the interior points 1..size-2 per axis.
#define STAGGERED_TENSOR_ITERATE_OVER_ALL_POINTS(tensor, include_border, CODE) \
  {                                                                            \
    const std::array<size_t, 3> &mif_sz_ = (tensor).sizes();                   \
    constexpr size_t mif_lo_ = (include_border) ? 0 : 1;                       \
    constexpr size_t mif_cut_ = (include_border) ? 0 : 1;                      \
    for (size_t k = mif_lo_; k < mif_sz_[2] - mif_cut_; k++)                   \
      for (size_t j = mif_lo_; j < mif_sz_[1] - mif_cut_; j++)                 \
        for (size_t i = mif_lo_; i < mif_sz_[0] - mif_cut_; i++) {             \
          CODE                                                                 \
        }                                                                      \
  }

// The owner points: periodic ghosts excluded, wall points included.
#define STAGGERED_TENSOR_ITERATE_OVER_ALL_OWNER_POINTS(tensor, CODE)               \
  {                                                                                \
    const std::array<size_t, 3> &mif_sz_ = (tensor).sizes();                       \
    const auto &mif_pbc_ = (tensor).constants.periodic_bc;                         \
    const size_t mif_lo_[3] = {mif_pbc_[0] ? 1u : 0u, mif_pbc_[1] ? 1u : 0u,       \
                               mif_pbc_[2] ? 1u : 0u};                             \
    const size_t mif_hi_[3] = {mif_sz_[0] - mif_lo_[0], mif_sz_[1] - mif_lo_[1],   \
                               mif_sz_[2] - mif_lo_[2]};                           \
    for (size_t k = mif_lo_[2]; k < mif_hi_[2]; k++)                               \
      for (size_t j = mif_lo_[1]; j < mif_hi_[1]; j++)                             \
        for (size_t i = mif_lo_[0]; i < mif_hi_[0]; i++) {                         \
          CODE                                                                     \
        }                                                                          \
  }

// tensor(i, j, k) = function(args...) over the points selected above.
#define STAGGERED_TENSOR_FUNCTION_ON_ALL_POINTS(tensor, function, include_border, args...) \
  STAGGERED_TENSOR_ITERATE_OVER_ALL_POINTS(tensor, include_border, (tensor)(i, j, k) = function(args);)

#define VELOCITY_TENSOR_SET_FOR_ALL_POINTS(velocity, f_u, f_v, f_w, include_border, args...) \
  {                                                                                          \
    STAGGERED_TENSOR_FUNCTION_ON_ALL_POINTS((velocity).u, f_u, include_border, args)         \
    STAGGERED_TENSOR_FUNCTION_ON_ALL_POINTS((velocity).v, f_v, include_border, args)         \
    STAGGERED_TENSOR_FUNCTION_ON_ALL_POINTS((velocity).w, f_w, include_border, args)         \
  }

#endif  // MIF_SHIM_STAGGERED_TENSOR_MACROS_H
