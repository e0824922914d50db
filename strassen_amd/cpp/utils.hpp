// utils.hpp -- helpers of the reference's utils.hpp:6-32 used by callers.
#ifndef SA_UTILS_HPP
#define SA_UTILS_HPP
#include <cstddef>

template <typename I1, typename I2>
constexpr I1 integer_division_round_up(const I1& lhs, const I2& rhs) {
  return lhs == 0 ? 0 : 1 + (lhs - 1) / rhs;
}
// Fortran-order 3D / 2D indices (utils.hpp:24-32)
constexpr size_t ijk(size_t i, size_t j, size_t k, size_t dimi, size_t dimj) { return i + j * dimi + k * dimi * dimj; }
constexpr size_t ij(size_t i, size_t j, size_t dimi) { return i + j * dimi; }
#endif
