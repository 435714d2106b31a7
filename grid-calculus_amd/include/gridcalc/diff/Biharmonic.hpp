// diff::biharmonic<Order> / diff::d4 (diff/Biharmonic.hpp:48-119 of the reference) on the
// device: the direct stencil sum_a d_a^4 + 2 sum_{a<b} d_a^2 d_b^2 (gc_biharmonic).
#pragma once

#include <gridcalc/core/Field.hpp>
#include <gridcalc/diff/detail/PreconditionAxisExtent.hpp>
#include <gridcalc/hip/Runtime.hpp>
#include <gridcalc/stencil/CentralDifference.hpp>
#include <gridcalc/stencil/FourthDerivative.hpp>

namespace gridcalc::diff {

namespace detail {
template <class T>
inline core::Field<T> biharmonicImpl(const core::Field<T>& psi, int order, int radius) {
  const auto& grid = psi.getGrid();
  requireAxisExtent("x", grid.getNx(), radius);
  requireAxisExtent("y", grid.getNy(), radius);
  requireAxisExtent("z", grid.getNz(), radius);
  core::Field<T> result(grid, core::detail::device_only);
  const gc_grid g = hip::toGrid(grid);
  hip::check(gc_biharmonic(hip::context(), &g, hip::dtype_of<T>::value, order, psi.deviceData(),
                           result.deviceDataDiscard()));
  return result;
}
}  // namespace detail

template <int Order = 2>
inline core::Field<double> biharmonic(const core::Field<double>& psi) {
  using S2 = stencil::Coefficients<Order>;
  using S4 = stencil::FourthDerivative<Order>;
  constexpr int max_radius = (S4::radius > S2::radius) ? S4::radius : S2::radius;
  return detail::biharmonicImpl(psi, Order, max_radius);
}

template <int Order = 2>
inline core::Field<float> biharmonic(const core::Field<float>& psi) {
  using S2 = stencil::Coefficients<Order>;
  using S4 = stencil::FourthDerivative<Order>;
  constexpr int max_radius = (S4::radius > S2::radius) ? S4::radius : S2::radius;
  return detail::biharmonicImpl(psi, Order, max_radius);
}

template <int Order = 2>
inline core::Field<double> d4(const core::Field<double>& psi) {
  return biharmonic<Order>(psi);
}

}  // namespace gridcalc::diff
