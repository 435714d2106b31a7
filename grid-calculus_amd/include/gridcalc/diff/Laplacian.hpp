// diff::laplacian<Order> (diff/Laplacian.hpp:44-75 of the reference) on the device:
// per-axis sums in stencil-offset order, then dxx/hx^2 + dyy/hy^2 + dzz/hz^2, computed by
// gc_laplacian.  The double overload is bit-identical to the reference arithmetic; the
// float overload is this library's single-precision form.
#pragma once

#include <gridcalc/core/Field.hpp>
#include <gridcalc/diff/detail/PreconditionAxisExtent.hpp>
#include <gridcalc/hip/Runtime.hpp>
#include <gridcalc/stencil/CentralDifference.hpp>

namespace gridcalc::diff {

namespace detail {
template <class T>
inline core::Field<T> laplacianImpl(const core::Field<T>& field, int order, int radius) {
  const auto& grid = field.getGrid();
  requireAxisExtent("x", grid.getNx(), radius);
  requireAxisExtent("y", grid.getNy(), radius);
  requireAxisExtent("z", grid.getNz(), radius);
  core::Field<T> result(grid, core::detail::device_only);
  const gc_grid g = hip::toGrid(grid);
  hip::check(gc_laplacian(hip::context(), &g, hip::dtype_of<T>::value, order, field.deviceData(),
                          result.deviceDataDiscard()));
  return result;
}
}  // namespace detail

template <int Order = 2>
inline core::Field<double> laplacian(const core::Field<double>& field) {
  using Coeffs = stencil::Coefficients<Order>;
  return detail::laplacianImpl(field, Order, Coeffs::radius);
}

template <int Order = 2>
inline core::Field<float> laplacian(const core::Field<float>& field) {
  using Coeffs = stencil::Coefficients<Order>;
  return detail::laplacianImpl(field, Order, Coeffs::radius);
}

}  // namespace gridcalc::diff
