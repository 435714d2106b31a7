// diff::divergence<Order> (diff/Divergence.hpp:43-135 of the reference) on the device:
// Vec3 -> scalar (gc_divergence) and Mat3 -> Vec3, row-wise d_j M(i, j) (gc_divergence_mat).
#pragma once

#include <gridcalc/diff/Gradient.hpp>

namespace gridcalc::diff {

template <int Order = 2>
inline core::Field<double> divergence(const core::Field<core::Vec3d>& vfield) {
  using Coeffs = stencil::FirstDerivative<Order>;
  return detail::firstDerivOp<double>(vfield, Order, Coeffs::radius, &gc_divergence);
}

template <int Order = 2>
inline core::Field<core::Vec3d> divergence(const core::Field<core::Mat3d>& mfield) {
  using Coeffs = stencil::FirstDerivative<Order>;
  return detail::firstDerivOp<core::Vec3d>(mfield, Order, Coeffs::radius, &gc_divergence_mat);
}

}  // namespace gridcalc::diff
