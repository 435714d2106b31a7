// The named second-order partials (diff/MixedPartial.hpp of the reference): each is
// detail::mixedDerivative<Nx, Ny, Nz, Order> on the device (gc_mixed_partial), Order = 2 by
// default as in the reference, for Field<double> (bit-identical) and Field<float>.
#pragma once

#include <gridcalc/core/Field.hpp>
#include <gridcalc/diff/detail/MultiIndexDerivative.hpp>

namespace gridcalc::diff {

template <int Order = 2, class T>
inline core::Field<T> dxx(const core::Field<T>& psi) {
  return detail::mixedDerivative<2, 0, 0, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> dyy(const core::Field<T>& psi) {
  return detail::mixedDerivative<0, 2, 0, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> dzz(const core::Field<T>& psi) {
  return detail::mixedDerivative<0, 0, 2, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> dxy(const core::Field<T>& psi) {
  return detail::mixedDerivative<1, 1, 0, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> dxz(const core::Field<T>& psi) {
  return detail::mixedDerivative<1, 0, 1, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> dyz(const core::Field<T>& psi) {
  return detail::mixedDerivative<0, 1, 1, Order>(psi);
}

}  // namespace gridcalc::diff
