// First-derivative central-difference weights (stencil/FirstDerivative.hpp:44-86 of the
// reference): unscaled; the consumer multiplies by 1/h.
#pragma once

#include <array>

namespace gridcalc::stencil {

template <int Order>
struct FirstDerivative;

template <>
struct FirstDerivative<2> {
  static constexpr int radius = 1;
  static constexpr std::array<int, 3> offsets = {-1, 0, 1};
  static constexpr std::array<double, 3> weights = {-0.5, 0.0, 0.5};
};

template <>
struct FirstDerivative<4> {
  static constexpr int radius = 2;
  static constexpr std::array<int, 5> offsets = {-2, -1, 0, 1, 2};
  static constexpr std::array<double, 5> weights = {1.0 / 12.0, -2.0 / 3.0, 0.0, 2.0 / 3.0,
                                                    -1.0 / 12.0};
};

}  // namespace gridcalc::stencil
