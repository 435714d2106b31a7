// Third-derivative central-difference weights (stencil/ThirdDerivative.hpp:38-84 of the
// reference): unscaled; the consumer multiplies by 1/h^3.
#pragma once

#include <array>

namespace gridcalc::stencil {

template <int Order>
struct ThirdDerivative;

template <>
struct ThirdDerivative<2> {
  static constexpr int radius = 2;
  static constexpr std::array<int, 5> offsets = {-2, -1, 0, 1, 2};
  static constexpr std::array<double, 5> weights = {-1.0 / 2.0, 1.0, 0.0, -1.0, 1.0 / 2.0};
};

template <>
struct ThirdDerivative<4> {
  static constexpr int radius = 3;
  static constexpr std::array<int, 7> offsets = {-3, -2, -1, 0, 1, 2, 3};
  static constexpr std::array<double, 7> weights = {1.0 / 8.0,   -1.0, 13.0 / 8.0, 0.0,
                                                    -13.0 / 8.0, 1.0,  -1.0 / 8.0};
};

}  // namespace gridcalc::stencil
