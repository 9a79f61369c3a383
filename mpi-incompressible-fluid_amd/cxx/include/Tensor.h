// A dense host array with x-fastest indexing (the storage convention of the
// reference's Tensor<Type, SpaceDim, DimensionsType>, include/Tensor.h:68,
// 219-238): element (i, j, k) sits at i + n0 * (j + n1 * k).
#ifndef MIF_SHIM_TENSOR_H
#define MIF_SHIM_TENSOR_H

#include <array>
#include <cstddef>
#include <numeric>
#include <utility>
#include <vector>

namespace mif {

template <typename Type, unsigned SpaceDim, typename DimensionsType>
class Tensor {
 public:
  explicit Tensor(const std::array<DimensionsType, SpaceDim> &dimensions)
      : dims_(dimensions), data_(count(dimensions), Type(0)) {}
  Tensor(const Tensor &) = delete;
  Tensor &operator=(const Tensor &) = delete;

  const std::array<DimensionsType, SpaceDim> &sizes() const { return dims_; }
  size_t size() const { return data_.size(); }
  void *raw_data() { return data_.data(); }
  const void *raw_data() const { return data_.data(); }
  void swap_data(Tensor &other) { data_.swap(other.data_); }

  template <typename... Idx>
  Type &operator()(Idx... idx) {
    static_assert(sizeof...(Idx) == SpaceDim, "one index per dimension");
    return data_[flat(idx...)];
  }
  template <typename... Idx>
  const Type &operator()(Idx... idx) const {
    static_assert(sizeof...(Idx) == SpaceDim, "one index per dimension");
    return data_[flat(idx...)];
  }

 private:
  static size_t count(const std::array<DimensionsType, SpaceDim> &d) {
    size_t n = 1;
    for (auto v : d) n *= static_cast<size_t>(v);
    return n;
  }
  template <typename... Idx>
  size_t flat(Idx... idx) const {
    const size_t ix[] = {static_cast<size_t>(idx)...};
    size_t q = 0;
    for (int a = static_cast<int>(SpaceDim) - 1; a >= 0; a--) q = q * static_cast<size_t>(dims_[a]) + ix[a];
    return q;
  }

  std::array<DimensionsType, SpaceDim> dims_;
  std::vector<Type> data_;
};

}  // namespace mif

#endif  // MIF_SHIM_TENSOR_H
