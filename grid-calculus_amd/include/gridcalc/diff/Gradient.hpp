// diff::gradient<Order> (diff/Gradient.hpp:43-137 of the reference) on the device:
// scalar -> Vec3 (gc_gradient) and Vec3 -> Mat3 with M(i, j) = d_j v_i (gc_gradient_vec).
#pragma once

#include <gridcalc/core/EigenAliases.hpp>
#include <gridcalc/core/Field.hpp>
#include <gridcalc/diff/detail/PreconditionAxisExtent.hpp>
#include <gridcalc/hip/Runtime.hpp>
#include <gridcalc/stencil/FirstDerivative.hpp>

namespace gridcalc::diff {

namespace detail {
template <class Out, class In>
inline core::Field<Out> firstDerivOp(const core::Field<In>& in, int order, int radius,
                                     gc_status (*fn)(gc_context*, const gc_grid*, gc_dtype, int,
                                                     const void*, void*)) {
  const auto& grid = in.getGrid();
  requireAxisExtent("x", grid.getNx(), radius);
  requireAxisExtent("y", grid.getNy(), radius);
  requireAxisExtent("z", grid.getNz(), radius);
  core::Field<Out> result(grid, core::detail::device_only);
  const gc_grid g = hip::toGrid(grid);
  hip::check(fn(hip::context(), &g, GC_F64, order, in.deviceData(), result.deviceDataDiscard()));
  return result;
}
}  // namespace detail

template <int Order = 2>
inline core::Field<core::Vec3d> gradient(const core::Field<double>& field) {
  using Coeffs = stencil::FirstDerivative<Order>;
  return detail::firstDerivOp<core::Vec3d>(field, Order, Coeffs::radius, &gc_gradient);
}

template <int Order = 2>
inline core::Field<core::Mat3d> gradient(const core::Field<core::Vec3d>& field) {
  using Coeffs = stencil::FirstDerivative<Order>;
  return detail::firstDerivOp<core::Mat3d>(field, Order, Coeffs::radius, &gc_gradient_vec);
}

}  // namespace gridcalc::diff
