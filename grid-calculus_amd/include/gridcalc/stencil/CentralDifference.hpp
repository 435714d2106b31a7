// Second-derivative central-difference weights (stencil/CentralDifference.hpp:34-74 of the
// reference): unscaled, offsets -r..r; the consumer multiplies by 1/h^2.  Orders 2 and 4;
// any other Order is a compile error (undefined primary template).
#pragma once

#include <array>

namespace gridcalc::stencil {

template <int Order>
struct Coefficients;

template <>
struct Coefficients<2> {
  static constexpr int radius = 1;
  static constexpr std::array<int, 3> offsets = {-1, 0, 1};
  static constexpr std::array<double, 3> weights = {1.0, -2.0, 1.0};
};

template <>
struct Coefficients<4> {
  static constexpr int radius = 2;
  static constexpr std::array<int, 5> offsets = {-2, -1, 0, 1, 2};
  static constexpr std::array<double, 5> weights = {-1.0 / 12.0, 4.0 / 3.0, -5.0 / 2.0,
                                                    4.0 / 3.0, -1.0 / 12.0};
};

}  // namespace gridcalc::stencil
