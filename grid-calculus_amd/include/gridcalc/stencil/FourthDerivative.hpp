// Fourth-derivative central-difference weights (stencil/FourthDerivative.hpp:39-83 of the
// reference): unscaled; the consumer multiplies by 1/h^4.
#pragma once

#include <array>

namespace gridcalc::stencil {

template <int Order>
struct FourthDerivative;

template <>
struct FourthDerivative<2> {
  static constexpr int radius = 2;
  static constexpr std::array<int, 5> offsets = {-2, -1, 0, 1, 2};
  static constexpr std::array<double, 5> weights = {1.0, -4.0, 6.0, -4.0, 1.0};
};

template <>
struct FourthDerivative<4> {
  static constexpr int radius = 3;
  static constexpr std::array<int, 7> offsets = {-3, -2, -1, 0, 1, 2, 3};
  static constexpr std::array<double, 7> weights = {
      -1.0 / 6.0, 2.0, -13.0 / 2.0, 28.0 / 3.0, -13.0 / 2.0, 2.0, -1.0 / 6.0};
};

}  // namespace gridcalc::stencil
